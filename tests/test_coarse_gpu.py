"""GPU parity of the two-level preconditioner (svg_coarse_*) against its CPU
oracle (oracle/coarse.py): the aggregates exactly, the three products to the
FP32 arithmetic of the device (1e-4 relative), the Galerkin condition on the
device's output, and the flexible GMRES with the two-level preconditioner
against the FP64 true residual."""
import ctypes as C

import numpy as np
import pytest

import oracle
from oracle import coarse as OC
from svfsiplus_amd import _lib, workload as W

pytestmark = pytest.mark.gpu


def _scaled(p, seed=1):
    ph = W.Physics()
    rng = np.random.default_rng(seed)
    Ao, Yo, An, Yn = W.dirichlet_state(p, ph)
    Ag, Yg = oracle.pici(ph, Ao, An, Yo, Yn)
    Yg = Yg + 0.3 * rng.standard_normal(Yg.shape)
    R, V = oracle.construct_fluid(p, ph, Ag, Yg)
    V2, R2, Wd = oracle.precond_diag(p, V, R)
    return V2, R2, Wd


def _dev(torch, a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def _prepare(torch, p, V2, target):
    from svfsiplus_amd.newton import DeviceLhs, _stream
    L = DeviceLhs(p)
    dV = _dev(torch, V2)
    _lib.call("svg_coarse_prepare", L.h, target, dV.data_ptr(), _stream())
    na = C.c_int()
    agg = np.empty(p.nNo, dtype=np.int32)
    _lib.call("svg_coarse_info", L.h, C.byref(na), agg.ctypes.data)
    return L, dV, agg, na.value


@pytest.mark.parametrize("case", ["pipe4-t16", "pipe8-t32", "sten8-t64", "bif-t24"])
def test_aggregates_and_products_match_oracle(torch_cuda, case):
    from svfsiplus_amd.newton import _stream
    torch = torch_cuda
    if case.startswith("bif"):
        from svfsiplus_amd import batch as B
        p = B.make_batch(W.bifurcation_batch(2, tets_min=3e3, tets_max=6e3))[0]
    else:
        n = int(case.split("-")[0][4:]) if case.startswith("pipe") else 8
        p = W.make_pipe(n, sten_depth=0.5, sten_z0=1.25, sten_len=0.5) if case.startswith("sten") else W.make_pipe(n)
    target = int(case.split("-t")[1])
    V2, R2, Wd = _scaled(p, seed=len(case))
    L, dV, agg, na = _prepare(torch, p, V2, target)
    agg0, na0 = OC.aggregate(p.rowPtr, p.colPtr, target)
    assert na == na0 and np.array_equal(agg, agg0)
    T = OC.TwoLevel(p.rowPtr, p.colPtr, V2, agg0, na0)
    rng = np.random.default_rng(7)
    nvec = 3
    Vh = rng.standard_normal((nvec, p.nNo, 4)) * (Wd != 0)
    dVin = _dev(torch, Vh)
    for mode, ref in ((_lib.COARSE_APPLY, T.apply), (_lib.COARSE_RESIDUAL, T.residual), (_lib.COARSE_SMOOTH, T.smooth)):
        dZ = torch.empty_like(dVin)
        _lib.call("svg_coarse_apply", L.h, mode, nvec, dVin.data_ptr(), dZ.data_ptr(), _stream())
        Z = dZ.cpu().numpy()
        for b in range(nvec):
            z0 = ref(Vh[b].ravel())
            rel = np.linalg.norm(Z[b].ravel() - z0) / np.linalg.norm(z0)
            print(f"{case} ({na} aggregates) mode {mode} vec {b}: rel {rel:.2e}")
            assert rel <= 1e-4
            if mode == _lib.COARSE_RESIDUAL:
                # the Galerkin condition on the device's own output: zero aggregate pressure sums
                s = np.bincount(agg, weights=Z[b][:, 3], minlength=na)
                s0 = np.bincount(agg, weights=Vh[b][:, 3], minlength=na)
                assert np.linalg.norm(s) <= 1e-4 * np.linalg.norm(s0)
    L.free()


def test_errors_are_reported(torch_cuda):
    from svfsiplus_amd.newton import DeviceLhs, _stream
    torch = torch_cuda
    p = W.make_pipe(4)
    L = DeviceLhs(p)
    x = torch.zeros((p.nNo, 4), dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    lib = _lib.load()
    assert lib.svg_coarse_apply(L.h, 0, 1, x.data_ptr(), y.data_ptr(), _stream()) == _lib.SVG_ERR_ARG  # not prepared
    V2, _, _ = _scaled(p)
    L2, dV, _, _ = _prepare(torch, p, V2, 16)
    assert lib.svg_coarse_apply(L2.h, 7, 1, x.data_ptr(), y.data_ptr(), _stream()) == _lib.SVG_ERR_ARG  # mode
    assert lib.svg_coarse_prepare(L2.h, 1, dV.data_ptr(), _stream()) == _lib.SVG_ERR_ARG  # aggregates of one node
    # more aggregates than the dense coarse matrix allows: refused on the host, before any device work
    big = W.make_pipe(16)
    Lb = DeviceLhs(big)
    dVb = torch.zeros((big.nnz, 16), dtype=torch.float64, device="cuda")
    assert lib.svg_coarse_prepare(Lb.h, 2, dVb.data_ptr(), _stream()) == _lib.SVG_ERR_ARG
    assert b"8192" in lib.svg_last_error()
    for o in (L, L2, Lb):
        o.free()


@pytest.mark.parametrize("precision", [0, 1])
def test_two_level_gmres_solve(torch_cuda, precision):
    """The flexible GMRES with the two-level preconditioner (FP64, and the
    mixed-precision form) on the first Newton system of the n = 8 pipe: the
    FP64 true residual of the scaled system below 1e-6, a third of the plain
    solve's iterations or fewer, the solution within 1e-5 of the FP64 oracle's,
    and two runs bit for bit equal (a deterministic preconditioner)."""
    from svfsiplus_amd.newton import DeviceLhs, fsils_solve, ls_create
    torch = torch_cuda
    p = W.make_pipe(8)
    ph = W.Physics()
    Ao, Yo, An, Yn = W.dirichlet_state(p, ph)
    Ag, Yg = oracle.pici(ph, Ao, An, Yo, Yn)
    R, V = oracle.construct_fluid(p, ph, Ag, Yg)
    x0, V0, ols = oracle.fsils_solve(p, oracle.ls_create(oracle.LS_GMRES, relTol=1e-10, absTol=1e-30, maxItr=2000,
                                                         dimKry=50), R, V)
    V2, R2, Wd = oracle.precond_diag(p, V, R)
    out = []
    for rep in range(2):
        L = DeviceLhs(p)
        ls = ls_create(_lib.LS_GMRES, relTol=1e-6, absTol=1e-30, maxItr=1000, dimKry=50)
        dR, dV = _dev(torch, R), _dev(torch, V)
        fsils_solve(L, ls, 4, dR, dV, precision=precision, coarse=32)
        x = dR.cpu().numpy()
        # true residual of the scaled system: y = x / W on the free dofs
        y = np.divide(x, Wd, out=np.zeros_like(x), where=Wd != 0)
        res = R2 - oracle.spar_mul_vv(p, 4, V2, y)
        rel = np.linalg.norm(res) / np.linalg.norm(R2)
        err = np.linalg.norm(x - x0) / np.linalg.norm(x0)
        print(f"precision {precision}: itr {ls.RI.itr} (plain 1e-10: {ols.RI.itr}) true rel res {rel:.2e} err {err:.2e}")
        assert ls.RI.suc == 1 and rel <= 1e-6
        assert err <= 1e-5
        out.append((x, ls.RI.itr))
        L.free()
    assert out[0][1] == out[1][1] and np.array_equal(out[0][0], out[1][0])
    plain = ls_create(_lib.LS_GMRES, relTol=1e-6, absTol=1e-30, maxItr=1000, dimKry=50)
    L = DeviceLhs(p)
    dR, dV = _dev(torch, R), _dev(torch, V)
    fsils_solve(L, plain, 4, dR, dV, precision=precision)
    assert 3 * out[0][1] <= plain.RI.itr
    L.free()


@pytest.mark.parametrize("precision", [_lib.PRECISION_FP64, _lib.PRECISION_MIXED])
def test_two_level_keeps_coarse_inverse(torch_cuda, precision):
    """coarse_keep: a first solve with coarse_keep set but no inverse on the lhs
    builds it (bit for bit the solve without the flag); a second system on the
    same lhs (the next Newton iteration's, a perturbed state) solved with the
    kept inverse reaches the FP64 true residual 1e-6 of its own scaled system,
    and its iteration count stays within 25 % of the rebuilt one's."""
    from svfsiplus_amd.newton import DeviceLhs, fsils_solve, ls_create
    torch = torch_cuda
    p = W.make_pipe(8)
    ph = W.Physics()
    Ao, Yo, An, Yn = W.dirichlet_state(p, ph)
    Ag, Yg = oracle.pici(ph, Ao, An, Yo, Yn)
    R1, V1 = oracle.construct_fluid(p, ph, Ag, Yg)
    rng = np.random.default_rng(3)
    R2, V2 = oracle.construct_fluid(p, ph, Ag + rng.standard_normal(Ag.shape), Yg + 0.2 * rng.standard_normal(Yg.shape))
    mk = lambda: ls_create(_lib.LS_GMRES, relTol=1e-6, absTol=1e-30, maxItr=1000, dimKry=50)
    res = {}
    for keep in (False, True):
        L = DeviceLhs(p)
        out = []
        for R, V in ((R1, V1), (R2, V2)):
            ls = mk()
            dR, dV = _dev(torch, R), _dev(torch, V)
            fsils_solve(L, ls, 4, dR, dV, precision=precision, coarse=32, coarse_keep=keep)
            x = dR.cpu().numpy()
            Vs, Rs, Wd = oracle.precond_diag(p, V, R)
            y = np.divide(x, Wd, out=np.zeros_like(x), where=Wd != 0)
            rel = np.linalg.norm(Rs - oracle.spar_mul_vv(p, 4, Vs, y)) / np.linalg.norm(Rs)
            print(f"keep {keep}: itr {ls.RI.itr} true rel res {rel:.2e}")
            assert ls.RI.suc == 1 and rel <= 1e-6
            out.append((x, ls.RI.itr))
        res[keep] = out
        L.free()
    assert np.array_equal(res[True][0][0], res[False][0][0]) and res[True][0][1] == res[False][0][1]
    assert res[True][1][1] <= 1.25 * res[False][1][1] + 2


def test_timed_launches_are_the_solve_iterations(torch_cuda):
    """With every launch timed, a mixed-precision two-level solve keeps exactly
    its real work: the FP32 SpMV samples number its Arnoldi steps and the FP64
    SpMV samples its restart checks, together ls.itr. The launches that return
    at once because the solve is already done (the steps the host enqueues past
    the converging one, and the cycle after a restart check that met the
    tolerance) are dropped; a mistake in either would break the identity."""
    import ctypes as C
    from svfsiplus_amd.newton import DeviceLhs, fsils_solve, ls_create
    torch = torch_cuda
    p = W.make_pipe(8)
    ph = W.Physics()
    Ao, Yo, An, Yn = W.dirichlet_state(p, ph)
    Ag, Yg = oracle.pici(ph, Ao, An, Yo, Yn)
    R, V = oracle.construct_fluid(p, ph, Ag, Yg)
    lib = _lib.load()
    L = DeviceLhs(p)
    # a first solve tunes the SpMV launch variants (untimed)
    fsils_solve(L, ls_create(_lib.LS_GMRES, relTol=1e-6, absTol=1e-30, maxItr=1000, dimKry=50), 4, _dev(torch, R),
                _dev(torch, V), precision=_lib.PRECISION_MIXED, coarse=32)
    _lib.check(lib.svg_prof_sample(1))
    _lib.check(lib.svg_prof_only(None))
    _lib.check(lib.svg_prof_reset())
    _lib.check(lib.svg_prof_enable(1))
    try:
        ls = ls_create(_lib.LS_GMRES, relTol=1e-6, absTol=1e-30, maxItr=1000, dimKry=50)
        fsils_solve(L, ls, 4, _dev(torch, R), _dev(torch, V), precision=_lib.PRECISION_MIXED, coarse=32)
    finally:
        lib.svg_prof_enable(0)
    counts = {}
    for name in ("spmv32", "spmv_bcsr"):
        cnt, ms = C.c_int(), C.c_double()
        _lib.check(lib.svg_prof_get(name.encode(), C.byref(cnt), C.byref(ms)))
        counts[name] = cnt.value
    print(f"itr {ls.RI.itr}, timed launches {counts}")
    assert ls.RI.suc == 1
    assert counts["spmv_bcsr"] >= 2  # the restart checks, the last one finding the tolerance met
    assert counts["spmv32"] + counts["spmv_bcsr"] == ls.RI.itr
    L.free()
