"""GPU parity of the HIP path against the CPU oracle (same seeded inputs).

Bars: bit-exact for the block SpMV, the diagonal preconditioner and the
assembled Val and R (same arithmetic, element contributions added in element
order); GMRES iteration counts within +-1 of the oracle's and solutions within
the solver tolerance; Newton iteration records matching the oracle's.
"""
import numpy as np
import pytest

import oracle
import sys, os
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from svfsiplus_amd import _lib, workload as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pipe4():
    return W.make_pipe(4)


@pytest.fixture(scope="module")
def phys():
    return W.Physics()


def _dev(torch, a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def _state(pipe, phys, seed=1):
    rng = np.random.default_rng(seed)
    Ao, Yo, An, Yn = W.dirichlet_state(pipe, phys)
    Ag, Yg = oracle.pici(phys, Ao, An, Yo, Yn)
    Yg = Yg + 0.5 * rng.standard_normal(Yg.shape)
    Ag = Ag + 10.0 * rng.standard_normal(Ag.shape)
    return Ag, Yg


def test_spar_mul_vv_bitwise(torch_cuda, pipe4):
    from svfsiplus_amd.newton import DeviceLhs, _stream
    torch = torch_cuda
    rng = np.random.default_rng(7)
    L = DeviceLhs(pipe4)
    for dof in (4, 3, 2, 1):
        K = rng.standard_normal((pipe4.nnz, dof * dof))
        U = rng.standard_normal((pipe4.nNo, dof))
        ref = oracle.spar_mul_vv(pipe4, dof, K, U)
        dK, dU = _dev(torch, K), _dev(torch, U)
        dKU = torch.empty((pipe4.nNo, dof), dtype=torch.float64, device="cuda")
        _lib.call("svg_spar_mul_vv", L.h, dof, dK.data_ptr(), dU.data_ptr(), dKU.data_ptr(), _stream())
        torch.cuda.synchronize()
        got = dKU.cpu().numpy()
        print("spmv dof", dof, "max diff", np.abs(got - ref).max())
        assert np.array_equal(got, ref), f"dof={dof}"


@pytest.mark.parametrize("n", [4, 8])
def test_every_spmv_variant_bitwise(torch_cuda, n):
    """Every launch variant the FP64 SpMV autotuner may pick (unroll depth,
    non-temporal streaming, XCD row bands) gives the oracle's product bit for
    bit: the tuner chooses on time alone, so any of them may run the solve."""
    import ctypes as C
    from svfsiplus_amd.newton import DeviceLhs, _stream
    torch = torch_cuda
    pipe = W.make_pipe(n)
    rng = np.random.default_rng(11)
    L = DeviceLhs(pipe)
    K = rng.standard_normal((pipe.nnz, 16))
    U = rng.standard_normal((pipe.nNo, 4))
    ref = oracle.spar_mul_vv(pipe, 4, K, U)
    dK, dU = _dev(torch, K), _dev(torch, U)
    lib = _lib.load()
    nv = C.c_int()
    _lib.check(lib.svg_spmv_num_variants(C.byref(nv)))
    assert nv.value >= 12
    for v in range(nv.value):
        dKU = torch.full((pipe.nNo, 4), float("nan"), dtype=torch.float64, device="cuda")
        us = C.c_double()
        _lib.check(lib.svg_spmv_time_variant(L.h, v, dK.data_ptr(), dU.data_ptr(), dKU.data_ptr(), 1, C.byref(us),
                                             _stream()))
        torch.cuda.synchronize()
        got = dKU.cpu().numpy()
        assert np.array_equal(got, ref), f"variant {v}: max diff {np.nanmax(np.abs(got - ref))}"
    L.free()


def test_precond_diag_bitwise(torch_cuda, pipe4, phys):
    from svfsiplus_amd.newton import DeviceLhs, _stream
    torch = torch_cuda
    Ag, Yg = _state(pipe4, phys)
    R, V = oracle.construct_fluid(pipe4, phys, Ag, Yg)
    V2, R2, W2 = oracle.precond_diag(pipe4, V, R)
    L = DeviceLhs(pipe4)
    dV, dR = _dev(torch, V), _dev(torch, R)
    dW = torch.empty((pipe4.nNo, 4), dtype=torch.float64, device="cuda")
    _lib.call("svg_precond_diag", L.h, 4, dV.data_ptr(), dR.data_ptr(), dW.data_ptr(), _stream())
    torch.cuda.synchronize()
    assert np.array_equal(dW.cpu().numpy(), W2)
    assert np.array_equal(dR.cpu().numpy(), R2)
    assert np.array_equal(dV.cpu().numpy(), V2)


def test_dot_norm(torch_cuda, pipe4):
    from svfsiplus_amd.newton import _stream
    import ctypes as C
    torch = torch_cuda
    rng = np.random.default_rng(3)
    for n in (1, 63, 1000, 4 * pipe4.nNo, 300001):
        U = rng.standard_normal(n)
        V = rng.standard_normal(n)
        dU, dV = _dev(torch, U), _dev(torch, V)
        out = C.c_double()
        _lib.call("svg_dot_v", 1, n, dU.data_ptr(), dV.data_ptr(), C.byref(out), _stream())
        exact = float(np.dot(U.astype(np.longdouble), V.astype(np.longdouble)))
        assert abs(out.value - exact) <= 1e-14 * np.abs(U * V).sum() + 1e-300
        _lib.call("svg_norm_v", 1, n, dU.data_ptr(), C.byref(out), _stream())
        assert abs(out.value - np.sqrt(np.sum(U.astype(np.longdouble) ** 2))) <= 1e-14 * out.value
    # an empty vector (nNo = 0): the sums are 0, as the reference's loops give
    out = C.c_double(1.0)
    _lib.call("svg_dot_v", 4, 0, dU.data_ptr(), dV.data_ptr(), C.byref(out), _stream())
    assert out.value == 0.0
    out = C.c_double(1.0)
    _lib.call("svg_norm_v", 4, 0, dU.data_ptr(), C.byref(out), _stream())
    assert out.value == 0.0


@pytest.mark.parametrize("n", [2, 4, 8])
def test_construct_fluid(torch_cuda, phys, n):
    from svfsiplus_amd.newton import DeviceLhs, DeviceMesh, construct_fluid
    torch = torch_cuda
    pipe = W.make_pipe(n)
    Ag, Yg = _state(pipe, phys, seed=n)
    R0, V0 = oracle.construct_fluid(pipe, phys, Ag, Yg)
    L = DeviceLhs(pipe)
    M = DeviceMesh(L, pipe)
    dR = torch.zeros((pipe.nNo, 4), dtype=torch.float64, device="cuda")
    dV = torch.zeros((pipe.nnz, 16), dtype=torch.float64, device="cuda")
    construct_fluid(M, phys, _dev(torch, Ag), _dev(torch, Yg), dR, dV)
    torch.cuda.synchronize()
    R1, V1 = dR.cpu().numpy(), dV.cpu().numpy()
    print(f"n={n} R differs in {(R1 != R0).sum()} of {R0.size}, Val in {(V1 != V0).sum()} of {V0.size}")
    assert np.array_equal(R1, R0) and np.array_equal(V1, V0)


@pytest.mark.parametrize("n", [4, 8])
def test_gmres_solve(torch_cuda, phys, n):
    """gmres_v on the device: bitwise equal to the oracle run with the device's
    reduction order, and within the solver tolerance of the reference-order oracle."""
    from svfsiplus_amd.newton import DeviceLhs, fsils_solve, ls_create
    torch = torch_cuda
    pipe = W.make_pipe(n)
    Ao, Yo, An, Yn = W.dirichlet_state(pipe, phys)
    Ag, Yg = oracle.pici(phys, Ao, An, Yo, Yn)
    R, V = oracle.construct_fluid(pipe, phys, Ag, Yg)
    tol = 1e-6
    mk = lambda: oracle.ls_create(oracle.LS_GMRES, relTol=tol, absTol=1e-30, maxItr=1000, dimKry=50)
    x0, V0, ols = oracle.fsils_solve(pipe, mk(), R, V)
    with oracle.device_order():
        xd, Vd, dls = oracle.fsils_solve(pipe, mk(), R, V)
    L = DeviceLhs(pipe)
    ls = ls_create(_lib.LS_GMRES, relTol=tol, absTol=1e-30, maxItr=1000, dimKry=50)
    dR, dV = _dev(torch, R), _dev(torch, V)
    fsils_solve(L, ls, 4, dR, dV)
    x1 = dR.cpu().numpy()
    rel = np.linalg.norm(x1 - x0) / np.linalg.norm(x0)
    print(f"n={n} itr gpu={ls.RI.itr} oracle(device order)={dls.RI.itr} oracle(reference order)={ols.RI.itr} "
          f"rel-to-reference-order={rel:.2e} max|gpu-device-order|={np.abs(x1 - xd).max():.2e}")
    assert ls.RI.itr == dls.RI.itr and ls.RI.suc == dls.RI.suc == 1
    assert ls.RI.iNorm == dls.RI.iNorm and ls.RI.fNorm == dls.RI.fNorm
    assert np.array_equal(x1, xd)
    assert np.array_equal(dV.cpu().numpy(), V0)
    assert abs(ls.RI.itr - ols.RI.itr) <= max(1, 0.02 * ols.RI.itr)
    assert rel <= 10 * tol


def test_newton_time_step(torch_cuda, phys):
    """Three Newton iterations (pici, construct_fluid, fsils_solve, picc) against the oracle:
    identical linear iteration counts and residual norms to the device-order oracle (assembly
    sums differ only in element order), and within tolerance of the reference-order oracle."""
    from svfsiplus_amd.newton import FluidNewton, ls_create
    pipe = W.make_pipe(4)
    tol = 1e-6
    mk = lambda: oracle.ls_create(oracle.LS_GMRES, relTol=tol, absTol=1e-30, maxItr=1000, dimKry=50)
    outlet = W.outlet_faces(pipe)
    An0, Yn0, hist0 = oracle.newton_time_step(pipe, phys, mk, outlet=outlet)
    with oracle.device_order():
        _, _, histd = oracle.newton_time_step(pipe, phys, mk, outlet=outlet)
    fn = FluidNewton(pipe, phys, ls_create(_lib.LS_GMRES, relTol=tol, absTol=1e-30, maxItr=1000, dimKry=50))
    hist1 = fn.time_step()
    for h0, hd, h1 in zip(hist0, histd, hist1):
        print("oracle(ref order)", h0, "\noracle(dev order)", hd, "\ngpu              ", h1)
    assert len(hist0) == len(hist1) == 3
    for h0, hd, h1 in zip(hist0, histd, hist1):
        assert abs(hd["ls_itr"] - h1["ls_itr"]) <= 1
        assert abs(h0["ls_itr"] - h1["ls_itr"]) <= max(1, 0.02 * h0["ls_itr"])
        assert abs(h0["iNorm"] - h1["iNorm"]) <= 1e-8 * h0["iNorm"]
    Yn1 = fn.Yn.cpu().numpy()
    assert np.linalg.norm(Yn1 - Yn0) <= 1e-6 * np.linalg.norm(Yn0)


def test_picp_set_bc_dir_bitwise(torch_cuda, phys):
    """svg_picp and svg_set_bc_dir against the oracle, bit for bit (all rows and a direction mask)."""
    torch = torch_cuda
    lib = _lib.load()
    rng = np.random.default_rng(21)
    n = 1000
    Ao, Yo = rng.standard_normal((n, 4)), rng.standard_normal((n, 4))
    An0, Yn0 = oracle.picp(phys, Ao, Yo)
    dAo, dYo = _dev(torch, Ao), _dev(torch, Yo)
    dAn = torch.empty_like(dAo)
    dYn = torch.empty_like(dYo)
    _lib.check(lib.svg_picp(n, phys.gam, dAo.data_ptr(), dYo.data_ptr(), dAn.data_ptr(), dYn.data_ptr(), None))
    torch.cuda.synchronize()
    assert np.array_equal(dAn.cpu().numpy(), An0) and np.array_equal(dYn.cpu().numpy(), Yn0)
    gN = rng.choice(n, 300, replace=False).astype(np.int32)
    for s0, lDof, eDir in ((0, 3, 0), (0, 2, 0b101), (1, 2, 0b011), (0, 4, 0)):
        gxn = rng.standard_normal((gN.size, lDof))
        A, Y = An0.copy(), Yn0.copy()
        oracle.set_bc_dir(gN, gxn, A, Y, s=s0, eDir=eDir, dirA=0.25, dirY=-1.5)
        dA, dY = _dev(torch, An0), _dev(torch, Yn0)
        dg = torch.tensor(gN, device="cuda")
        dx = _dev(torch, gxn)
        _lib.check(lib.svg_set_bc_dir(gN.size, dg.data_ptr(), 4, s0, lDof, eDir, dx.data_ptr(), 0.25, -1.5,
                                      dA.data_ptr(), dY.data_ptr(), None))
        torch.cuda.synchronize()
        assert np.array_equal(dA.cpu().numpy(), A) and np.array_equal(dY.cpu().numpy(), Y), (s0, lDof, eDir)
    # rows past tDof are an error, not a write
    assert lib.svg_set_bc_dir(gN.size, dg.data_ptr(), 4, 2, 3, 0, dx.data_ptr(), 0.0, 1.0, dA.data_ptr(),
                              dY.data_ptr(), None) != 0


def test_two_time_steps(torch_cuda, phys):
    """Two time steps on the device (advance: Ao = An, Yo = Yn, picp and
    set_bc_dir between them) against the oracle's time loop."""
    from svfsiplus_amd.newton import FluidNewton, ls_create
    pipe = W.make_pipe(4)
    tol = 1e-6
    mk = lambda: oracle.ls_create(oracle.LS_GMRES, relTol=tol, absTol=1e-30, maxItr=1000, dimKry=50)
    outlet = W.outlet_faces(pipe)
    An0, Yn0, hists0 = oracle.time_steps(pipe, phys, mk, 2, outlet=outlet)
    fn = FluidNewton(pipe, phys, ls_create(_lib.LS_GMRES, relTol=tol, absTol=1e-30, maxItr=1000, dimKry=50))
    hists1 = fn.time_steps(2)
    assert len(hists1) == 2
    for step, (h0s, h1s) in enumerate(zip(hists0, hists1)):
        assert len(h0s) == len(h1s) == 3
        for h0, h1 in zip(h0s, h1s):
            print(f"step {step + 1}: oracle itr {h0['ls_itr']} iNorm {h0['iNorm']:.9e} r1 {h0['r1']:.6e} | "
                  f"gpu itr {h1['ls_itr']} iNorm {h1['iNorm']:.9e} r1 {h1['r1']:.6e}")
            assert abs(h0["ls_itr"] - h1["ls_itr"]) <= max(1, 0.02 * h0["ls_itr"])
            assert abs(h0["iNorm"] - h1["iNorm"]) <= 1e-8 * h0["iNorm"]
            assert abs(h0["r1"] - h1["r1"]) <= 1e-8 * h0["r1"]
    # the second step starts from the first one's solution: different residuals
    assert hists0[1][0]["iNorm"] != hists0[0][0]["iNorm"]
    Yn1 = fn.Yn.cpu().numpy()
    assert np.linalg.norm(Yn1 - Yn0) <= 1e-6 * np.linalg.norm(Yn0)
    assert np.linalg.norm(fn.An.cpu().numpy() - An0) <= 1e-6 * np.linalg.norm(An0)


def test_restart_continues_bitwise(torch_cuda, phys, tmp_path):
    """write_restart after time step 1, init_from_bin in a new driver, time step 2:
    the same Newton records and the same Yn/An, bit for bit, as two steps run
    without the restart (the record carries Yn, An and eq.iNorm)."""
    from svfsiplus_amd import restart as RS
    from svfsiplus_amd.newton import FluidNewton, ls_create
    pipe = W.make_pipe(4)
    mk = lambda: ls_create(_lib.LS_GMRES, relTol=1e-6, absTol=1e-30, maxItr=1000, dimKry=50)
    fa = FluidNewton(pipe, phys, mk())
    ha = fa.time_steps(2)
    fb = FluidNewton(pipe, phys, mk())
    fb.time_step()
    f = fb.save_restart(str(tmp_path / "stFile"))
    assert os.path.basename(f) == "stFile_001.bin" and os.path.samefile(RS.last(str(tmp_path / "stFile")), f)
    fc = FluidNewton(pipe, phys, mk())
    r = fc.load_restart(RS.last(str(tmp_path / "stFile")))
    assert r.cTS == 1 and fc.step == 1
    hc = fc.time_step()
    print("continuous", ha[1], "\nrestarted ", hc)
    assert [dict(h) for h in hc] == [dict(h) for h in ha[1]]
    assert torch_cuda.equal(fc.Yn, fa.Yn) and torch_cuda.equal(fc.An, fa.An)


@pytest.mark.parametrize("name", ["pipe_n2", "pipe_n4"])
def test_gpu_against_reference_golden(torch_cuda, name):
    """The HIP path against the reference's own outputs (fixtures written from the reference build)."""
    import _golden as G
    from svfsiplus_amd.newton import DeviceLhs, DeviceMesh, construct_fluid, fsils_solve, ls_create
    torch = torch_cuda
    p, ph, d = G.load(name)
    L = DeviceLhs(p)
    M = DeviceMesh(L, p)
    dR = torch.zeros((p.nNo, 4), dtype=torch.float64, device="cuda")
    dV = torch.zeros((p.nnz, 16), dtype=torch.float64, device="cuda")
    construct_fluid(M, ph, _dev(torch, d["Ag_p"]), _dev(torch, d["Yg_p"]), dR, dV)
    torch.cuda.synchronize()
    # the reference's own construct_fluid output, bit for bit
    assert np.array_equal(dR.cpu().numpy(), d["R_p"]) and np.array_equal(dV.cpu().numpy(), d["Val_p"])
    tol = float(d["gm_tol"])
    ls = ls_create(_lib.LS_GMRES, relTol=tol, absTol=1e-30, maxItr=1000, dimKry=50)
    dR, dV = _dev(torch, d["R0"]), _dev(torch, d["Val0"])
    fsils_solve(L, ls, 4, dR, dV)
    itr, iNorm, fNorm, suc = d["gm_out"].tolist()
    x = dR.cpu().numpy()
    with oracle.device_order():
        xd, _, dls = oracle.fsils_solve(
            p, oracle.ls_create(oracle.LS_GMRES, relTol=tol, absTol=1e-30, maxItr=1000, dimKry=50), d["R0"], d["Val0"])
    print(name, "gmres itr gpu", ls.RI.itr, "reference", itr, "oracle(device order)", dls.RI.itr,
          "rel", np.linalg.norm(x - d["gm_x"]) / np.linalg.norm(d["gm_x"]))
    # classical Gram-Schmidt near Krylov exhaustion (tiny systems) amplifies the dot-product
    # rounding order, so the count is pinned to the device-order oracle, the solution to the reference
    assert ls.RI.itr == dls.RI.itr and ls.RI.suc == suc
    assert np.array_equal(x, xd)
    assert np.array_equal(dV.cpu().numpy(), d["gm_Val"])
    assert np.linalg.norm(x - d["gm_x"]) <= 10 * tol * np.linalg.norm(d["gm_x"])


def test_coupled_ns_solver(torch_cuda):
    """The NS solver with a coupled (resistance) outlet on the device: nS, add_bc_mul
    after the momentum products, op PRE in the inner GMRES and the Schur CG —
    bit for bit against the device-order oracle, and within the tolerance of the
    reference's own solution (golden coupled_n4, res = 100)."""
    import _golden as G
    from svfsiplus_amd.newton import DeviceLhs, fsils_solve, ls_create
    torch = torch_cuda
    p, ph, d = G.load("coupled_n4")
    faces = [("inlet", d["inlet"]), ("wall", d["wall"]), ("outlet", d["outlet"], "neu", d["outlet_val"])]
    res = d["ns_res"]
    tol = float(d["ns_tol"])
    L = DeviceLhs(p, faces)
    ls = ls_create(_lib.LS_NS, relTol=tol, absTol=1e-30, maxItr=1000, dimKry=50, ns_gm=(1000, 1e-2),
                   ns_cg=(1000, 1e-2))
    dR, dV = _dev(torch, d["R0"]), _dev(torch, d["Val0"])
    fsils_solve(L, ls, 4, dR, dV, res=res)
    itr, iNorm, fNorm, suc, gmi, cgi, resm, resc = d["ns_out"].tolist()
    x = dR.cpu().numpy()
    rel = np.linalg.norm(x - d["ns_x"]) / np.linalg.norm(d["ns_x"])
    with oracle.device_order():
        xd, _, ols = oracle.fsils_solve(
            p, oracle.ls_create(oracle.LS_NS, relTol=tol, absTol=1e-30, maxItr=1000, dimKry=50, relTolIn=(1e-2, 1e-2),
                                absTolIn=(1e-30, 1e-30), maxItrIn=(1000, 1000)), d["R0"], d["Val0"],
            faces=faces, res=res)
    print("coupled NS gpu RI/GM/CG", ls.RI.itr, ls.GM.itr, ls.CG.itr, "Resm/Resc", ls.Resm, ls.Resc,
          "| reference", itr, gmi, cgi, resm, resc, "| oracle(device order)", ols.RI.itr, ols.GM.itr, ols.CG.itr,
          "| rel to reference", rel, "max|gpu-device-order|", np.abs(x - xd).max())
    assert (ls.RI.itr, ls.GM.itr, ls.CG.itr, ls.Resm, ls.Resc) == (ols.RI.itr, ols.GM.itr, ols.CG.itr, ols.Resm, ols.Resc)
    assert np.array_equal(x, xd)
    assert ls.RI.suc == suc and abs(ls.RI.itr - itr) <= 1
    assert rel <= 1e-2


@pytest.mark.parametrize("name", ["pipe_n2", "pipe_n4"])
def test_ns_solver_against_reference_golden(torch_cuda, name):
    """Bipartitioned NS solver (ns_solver.cpp) on the device against the reference's record."""
    import _golden as G
    from svfsiplus_amd.newton import DeviceLhs, fsils_solve, ls_create
    torch = torch_cuda
    p, ph, d = G.load(name)
    tol = float(d["ns_tol"])
    L = DeviceLhs(p)
    ls = ls_create(_lib.LS_NS, relTol=tol, absTol=1e-30, maxItr=1000, dimKry=50, ns_gm=(1000, 1e-2),
                   ns_cg=(1000, 1e-2))
    dR, dV = _dev(torch, d["R0"]), _dev(torch, d["Val0"])
    fsils_solve(L, ls, 4, dR, dV)
    itr, iNorm, fNorm, suc, gmi, cgi, resm, resc = d["ns_out"].tolist()
    x = dR.cpu().numpy()
    rel = np.linalg.norm(x - d["ns_x"]) / np.linalg.norm(d["ns_x"])
    with oracle.device_order():
        xd, _, ols = oracle.fsils_solve(
            p, oracle.ls_create(oracle.LS_NS, relTol=tol, absTol=1e-30, maxItr=1000, dimKry=50, relTolIn=(1e-2, 1e-2),
                                absTolIn=(1e-30, 1e-30), maxItrIn=(1000, 1000)), d["R0"], d["Val0"])
    print(name, "NS gpu RI/GM/CG", ls.RI.itr, ls.GM.itr, ls.CG.itr, "Resm/Resc", ls.Resm, ls.Resc,
          "| reference", itr, gmi, cgi, resm, resc, "| oracle(device order)", ols.RI.itr, ols.GM.itr, ols.CG.itr,
          ols.Resm, ols.Resc, "| rel to reference", rel, "max|gpu-device-order|", np.abs(x - xd).max())
    assert (ls.RI.itr, ls.GM.itr, ls.CG.itr, ls.Resm, ls.Resc) == (ols.RI.itr, ols.GM.itr, ols.CG.itr, ols.Resm, ols.Resc)
    assert np.array_equal(x, xd)
    assert ls.RI.suc == suc and abs(ls.RI.itr - itr) <= 1
    assert abs(ls.RI.iNorm - iNorm) <= 1e-12 * iNorm
    assert rel <= 1e-3


@pytest.mark.parametrize("n", [4, 8])
def test_mixed_precision_solve(torch_cuda, phys, n):
    """FP32-SpMV / FP64-refined GMRES: the FP64 true residual of the scaled system is below
    the tolerance, and the solution is within the tolerance of the FP64 oracle's."""
    from svfsiplus_amd.newton import DeviceLhs, fsils_solve, ls_create
    torch = torch_cuda
    pipe = W.make_pipe(n)
    Ao, Yo, An, Yn = W.dirichlet_state(pipe, phys)
    Ag, Yg = oracle.pici(phys, Ao, An, Yo, Yn)
    R, V = oracle.construct_fluid(pipe, phys, Ag, Yg)
    tol = 1e-6
    x0, Vs, ols = oracle.fsils_solve(pipe, oracle.ls_create(oracle.LS_GMRES, relTol=tol, absTol=1e-30, maxItr=1000,
                                                            dimKry=50), R, V)
    _, Rs, Wd = oracle.precond_diag(pipe, V, R)
    L = DeviceLhs(pipe)
    ls = ls_create(_lib.LS_GMRES, relTol=tol, absTol=1e-30, maxItr=1000, dimKry=50)
    dR, dV = _dev(torch, R), _dev(torch, V)
    fsils_solve(L, ls, 4, dR, dV, precision=_lib.PRECISION_MIXED)
    x1 = dR.cpu().numpy()
    mask = Wd != 0
    # unscaled back to the scaled system's unknown (0 on the Dirichlet rows, where W = 0)
    y = np.divide(x1, Wd, out=np.zeros_like(x1), where=mask)
    res = Rs - oracle.spar_mul_vv(pipe, 4, Vs, y)
    rel_res = np.linalg.norm(res) / np.linalg.norm(Rs)
    err = np.linalg.norm(x1 - x0) / np.linalg.norm(x0)
    print(f"n={n} mixed: itr {ls.RI.itr} (fp64 oracle {ols.RI.itr}) true rel residual {rel_res:.2e} "
          f"|x-x_fp64|/|x_fp64| {err:.2e}")
    assert ls.RI.suc == 1 and rel_res <= tol
    assert err <= 20 * tol


def test_cpp_driver_matches_python_driver(torch_cuda, tmp_path):
    """The C++ host driver (svg_newton_step, the reference's time loop over
    the C++ host API) and the Python driver give the same Newton records over
    two time steps, and write byte-identical restart files."""
    import json
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(_lib.LIB_PATH), "svg_newton_step")
    cpp = str(tmp_path / "cpp")
    r = subprocess.run([exe, "4", "1e-6", "50", "1000", "2", cpp], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    recs = [json.loads(line) for line in r.stdout.splitlines()[1:]]
    from svfsiplus_amd.newton import FluidNewton, ls_create
    pipe = W.make_pipe(4)
    ls = ls_create(_lib.LS_GMRES, relTol=1e-6, absTol=1e-30, maxItr=1000, dimKry=50)
    fn = FluidNewton(pipe, W.Physics(), ls)
    hist = [h for step in fn.time_steps(2) for h in step]
    py = fn.save_restart(str(tmp_path / "py"))
    assert os.path.exists(cpp + "_001.bin") and os.path.samefile(cpp + "_last.bin", cpp + "_002.bin")
    assert open(cpp + "_002.bin", "rb").read() == open(py, "rb").read()
    assert [x["step"] for x in recs] == [1, 1, 1, 2, 2, 2]
    print("C++ driver", [(x["itr"], x["ls_itr"], x["iNorm"]) for x in recs])
    print("Python driver", [(x["itr"], x["ls_itr"], x["iNorm"]) for x in hist])
    assert len(recs) == len(hist)
    for a, b in zip(recs, hist):
        assert a["ls_itr"] == b["ls_itr"]
        assert a["iNorm"] == b["iNorm"] and a["fNorm"] == b["fNorm"]


def test_b_assem_neu_bc_matches_oracle(torch_cuda, pipe4, phys):
    """Outlet Neumann/backflow assembly (b_assem_neu_bc + b_fluid) on the device
    against the oracle, with velocities that flow back through part of the
    outlet and a nonzero Neumann value, accumulated onto an assembled system."""
    from svfsiplus_amd.newton import DeviceFace, DeviceLhs, DeviceMesh, b_assem_neu_bc, construct_fluid
    torch = torch_cuda
    Ag, Yg = _state(pipe4, phys)
    rng = np.random.default_rng(5)
    Yg[:, 2] = rng.uniform(-20.0, 20.0, pipe4.nNo)
    hg = rng.standard_normal(pipe4.nNo)
    fi, ge, inner = W.outlet_faces(pipe4)
    R0, V0 = oracle.construct_fluid(pipe4, phys, Ag, Yg)
    oracle.b_assem_neu_fluid(pipe4, phys, Yg, fi, inner, R0, V0, hg=hg)
    L = DeviceLhs(pipe4)
    M = DeviceMesh(L, pipe4)
    Fa = DeviceFace(M, fi, ge)
    dR = torch.zeros((pipe4.nNo, 4), dtype=torch.float64, device="cuda")
    dV = torch.zeros((pipe4.nnz, 16), dtype=torch.float64, device="cuda")
    dY = _dev(torch, Yg)
    construct_fluid(M, phys, _dev(torch, Ag), dY, dR, dV)
    b_assem_neu_bc(Fa, phys, dY, dR, dV, hg=_dev(torch, hg))
    torch.cuda.synchronize()
    R1, V1 = dR.cpu().numpy(), dV.cpu().numpy()
    Rb, Vb = oracle.construct_fluid(pipe4, phys, Ag, Yg)
    dRo, dRg = R0 - Rb, R1 - Rb
    print(f"face R contribution |oracle| {np.linalg.norm(dRo):.3e} |gpu-oracle| {np.linalg.norm(dRg - dRo):.2e}")
    assert np.linalg.norm(dRo) > 0 and np.linalg.norm(V0 - Vb) > 0
    assert np.linalg.norm(R1 - R0) <= 1e-12 * np.linalg.norm(R0)
    assert np.linalg.norm(V1 - V0) <= 1e-12 * np.linalg.norm(V0)


def test_b_assem_neu_bc_against_reference_golden(torch_cuda):
    """The device's outlet Neumann/backflow assembly against what the
    reference's own b_assem_neu_bc (gnnb, b_fluid, do_assem) gave on the same
    mesh, velocity (flowing back through part of the outlet) and Neumann value
    (tests/golden/backflow_n4.npz): norm-wise and entry-wise to 1e-12."""
    import _golden as G
    from svfsiplus_amd.newton import DeviceFace, DeviceLhs, DeviceMesh, b_assem_neu_bc
    torch = torch_cuda
    p, ph, d = G.load("backflow_n4")
    ph.beta, ph.qmTRI3 = float(d["beta"]), float(d["qmTRI3"])
    L = DeviceLhs(p)
    M = DeviceMesh(L, p)
    Fa = DeviceFace(M, d["faceIEN"], d["gE"])
    dR = torch.zeros((p.nNo, 4), dtype=torch.float64, device="cuda")
    dV = torch.zeros((p.nnz, 16), dtype=torch.float64, device="cuda")
    b_assem_neu_bc(Fa, ph, _dev(torch, d["Yg"]), dR, dV, hg=_dev(torch, d["hg"]))
    torch.cuda.synchronize()
    R1, V1 = dR.cpu().numpy(), dV.cpu().numpy()
    eR = np.linalg.norm(R1 - d["R"]) / np.linalg.norm(d["R"])
    eV = np.linalg.norm(V1 - d["Val"]) / np.linalg.norm(d["Val"])
    print(f"backflow vs reference golden: R rel {eR:.2e} (max {np.abs(R1 - d['R']).max():.2e}), "
          f"Val rel {eV:.2e} (max {np.abs(V1 - d['Val']).max():.2e})")
    assert eR <= 1e-12 and eV <= 1e-12
    assert np.abs(R1 - d["R"]).max() <= 1e-12 * np.abs(d["R"]).max()
    assert np.abs(V1 - d["Val"]).max() <= 1e-12 * np.abs(d["Val"]).max()
    for o in (Fa, M, L):
        o.free()



def test_spmv32_matches_fp64_product(torch_cuda, pipe4):
    """The FP32 block SpMV of the mixed-precision solve against the FP64 product
    of the same FP32 inputs (FP32 accumulation: relative error ~1e-7)."""
    from svfsiplus_amd.newton import DeviceLhs
    torch = torch_cuda
    rng = np.random.default_rng(3)
    K = rng.standard_normal((pipe4.nnz, 16)).astype(np.float32)
    U = rng.standard_normal((pipe4.nNo, 4)).astype(np.float32)
    L = DeviceLhs(pipe4)
    dK, dU = torch.tensor(K, device="cuda"), torch.tensor(U, device="cuda")
    dKU = torch.zeros((pipe4.nNo, 4), dtype=torch.float64, device="cuda")
    _lib.check(_lib.load().svg_spmv32(L.h, dK.data_ptr(), dU.data_ptr(), dKU.data_ptr(), None))
    torch.cuda.synchronize()
    ref = oracle.spar_mul_vv(pipe4, 4, K.astype(np.float64), U.astype(np.float64))
    rel = np.linalg.norm(dKU.cpu().numpy() - ref) / np.linalg.norm(ref)
    print(f"spmv32 rel err vs FP64 product {rel:.2e}")
    assert rel < 1e-6


def _coupled_residual(p, d, x):
    """Relative residual of K x + res v (v^T x) = R on the rows that are not
    Dirichlet (dense product over the unscaled system)."""
    n = p.nNo
    K = np.zeros((4 * n, 4 * n))
    V0 = d["Val0"].reshape(-1, 4, 4)
    for a in range(n):
        for j in range(p.rowPtr[a], p.rowPtr[a + 1]):
            K[4 * a:4 * a + 4, 4 * p.colPtr[j]:4 * p.colPtr[j] + 4] = V0[j]
    fixed = np.zeros(4 * n, bool)
    for nodes in (d["inlet"], d["wall"]):
        for i in range(3):
            fixed[4 * nodes + i] = True
    v = np.zeros(4 * n)
    for a, node in enumerate(d["outlet"]):
        v[4 * node:4 * node + 3] = d["outlet_val"][a]
    v[fixed] = 0.0
    R = d["R0"].ravel()
    x = x.ravel()
    r = R - K @ x - d["res"][2] * v * (v @ x)
    return np.linalg.norm(r[~fixed]) / np.linalg.norm(R[~fixed])


def test_coupled_gmres(torch_cuda):
    """GMRES with a coupled (resistance) outlet, add_bc_mul after every block
    product: bit for bit the device-order oracle (itself equal to the
    reference's fsils_solve in reference order), within the solver tolerance
    of the reference's own solution; the mixed-precision and learned forms
    solve the same coupled system."""
    import _golden as G
    from svfsiplus_amd.lpc import LearnedPC
    from svfsiplus_amd.newton import DeviceLhs, fsils_solve, ls_create
    torch = torch_cuda
    p, ph, d = G.load("coupled_n4")
    faces = [("inlet", d["inlet"]), ("wall", d["wall"]), ("outlet", d["outlet"], "neu", d["outlet_val"])]
    res = d["res"]
    tol = float(d["gm_tol"])
    L = DeviceLhs(p, faces)
    ls = ls_create(_lib.LS_GMRES, relTol=tol, absTol=1e-30, maxItr=1000, dimKry=50)
    dR, dV = _dev(torch, d["R0"]), _dev(torch, d["Val0"])
    fsils_solve(L, ls, 4, dR, dV, res=res)
    x = dR.cpu().numpy()
    with oracle.device_order():
        xd, _, dls = oracle.fsils_solve(
            p, oracle.ls_create(oracle.LS_GMRES, relTol=tol, absTol=1e-30, maxItr=1000, dimKry=50), d["R0"],
            d["Val0"], faces=faces, res=res)
    itr, iNorm, fNorm, suc = d["gm_out"].tolist()
    print("coupled gmres itr gpu", ls.RI.itr, "oracle(device order)", dls.RI.itr, "reference", itr,
          "rel", np.linalg.norm(x - d["gm_x"]) / np.linalg.norm(d["gm_x"]), "residual", _coupled_residual(p, d, x))
    # as for the uncoupled golden solve: on this 900-unknown system GMRES(50) runs far past Krylov
    # exhaustion, where classical Gram-Schmidt amplifies the dot products' rounding order (the
    # reference's serial order needs 550 iterations, the device's tree order 211), so the count is
    # pinned to the device-order oracle and the solution to the reference
    assert (ls.RI.itr, ls.RI.iNorm, ls.RI.fNorm) == (dls.RI.itr, dls.RI.iNorm, dls.RI.fNorm)
    assert np.array_equal(x, xd)
    assert ls.RI.suc == suc == 1
    assert np.linalg.norm(x - d["gm_x"]) <= 10 * tol * np.linalg.norm(d["gm_x"])
    ref_res = _coupled_residual(p, d, d["gm_x"])
    net = LearnedPC.load(os.path.join(os.path.dirname(_lib.LIB_PATH), "data", "lpc_weights.npz"))
    for label, kw in (("mixed", dict(precision=_lib.PRECISION_MIXED)),
                      ("learned", dict(precision=_lib.PRECISION_MIXED, lpc=net, fallback_window=20, net_period=5))):
        ls2 = ls_create(_lib.LS_GMRES, relTol=tol, absTol=1e-30, maxItr=1000, dimKry=50)
        dR, dV = _dev(torch, d["R0"]), _dev(torch, d["Val0"])
        fsils_solve(L, ls2, 4, dR, dV, res=res, **kw)
        r2 = _coupled_residual(p, d, dR.cpu().numpy())
        print(label, "itr", ls2.RI.itr, "residual", r2, "reference's", ref_res)
        assert ls2.RI.suc == 1
        assert r2 <= 10 * max(ref_res, 1e-9)


def test_newton_step_on_a_loaded_model(torch_cuda, tmp_path, phys):
    """The seeded pipe written as a SimVascular-style model (volume .vtu, face
    .vtp files) and loaded back runs the same Newton step, record for record."""
    from svfsiplus_amd import vtk_io as V
    from svfsiplus_amd.newton import FluidNewton, ls_create
    p = W.make_pipe(4)
    V.write_vtu(str(tmp_path / "mesh.vtu"), p.x, p.IEN)
    paths = {}
    for name, bit in (("inlet", _lib.NODE_INLET), ("wall", _lib.NODE_WALL), ("outlet", _lib.NODE_OUTLET)):
        fi, ge, _ = W.outlet_faces(p, bit)
        nodes = np.unique(fi)
        V.write_vtp(str(tmp_path / f"{name}.vtp"), p.x[nodes], np.searchsorted(nodes, fi),
                    {"GlobalNodeID": (nodes + 1).astype(np.int32)}, {"GlobalElementID": (ge + 1).astype(np.int32)})
        paths[name] = str(tmp_path / f"{name}.vtp")
    q, faces = V.load_mesh(str(tmp_path / "mesh.vtu"), paths, dirichlet=["inlet", "wall"],
                           node_flags={"outlet": _lib.NODE_OUTLET})
    assert np.array_equal(q.faces[0][1], p.faces[0][1]) and np.array_equal(q.faces[1][1], p.faces[1][1])
    mk = lambda: ls_create(_lib.LS_GMRES, relTol=1e-6, absTol=1e-30, maxItr=1000, dimKry=50)
    h0 = FluidNewton(p, phys, mk()).time_step()
    fn = FluidNewton(q, phys, mk())
    h1 = fn.time_step()
    for a, b in zip(h0, h1):
        assert (a["ls_itr"], a["iNorm"], a["fNorm"]) == (b["ls_itr"], b["iNorm"], b["fNorm"])
    V.write_result(str(tmp_path / "result_001.vtu"), q, fn.Yn, fn.An)
    r = V.read(str(tmp_path / "result_001.vtu"))
    assert np.array_equal(r.point_data["Velocity"], fn.Yn.cpu().numpy()[:, :3])
