"""The drop-in back-end on the GPU, driven through the reference's own types and
Newton-iteration sequence (oracle/bind_harness.cpp: ls_alloc, construct_fluid
or the back-end's device assembly, b_assem_neu_bc, ls_solve with
GpuFsilsLinearAlgebra as the equation's LinearAlgebra):

* solver only (the reference's host assembly reaching the back-end through
  LinearAlgebra::assemble): bit for bit the C-ABI solve of the reference's own
  assembled system (golden R0 / Val0), and within the tolerance of the
  reference's own fsils_solve (golden gm_x / ns_x);
* device assembly (the branch at eq_assem.cpp:400-402): bit for bit the C-ABI
  construct_fluid (+ b_assem_neu_bc) and solve on the same inputs, and within
  the tolerance of the solver-only result;
* two Newton iterations of the same state give the same increment bit for bit
  (the device R / Val are cleared between iterations as ls_alloc clears the
  host ones)."""
import numpy as np
import pytest

import oracle
from oracle import bind as OB
from svfsiplus_amd import _lib

pytestmark = pytest.mark.gpu

# the settings of the golden solves (tests/golden/make_golden.py)
GM_PRM = oracle.ls_params(oracle.ls_create(oracle.LS_GMRES, relTol=1e-6, absTol=1e-30, maxItr=1000, dimKry=50))
NS_PRM = oracle.ls_params(oracle.ls_create(oracle.LS_NS, relTol=1e-3, absTol=1e-30, maxItr=1000, dimKry=50,
                                           relTolIn=(1e-2, 1e-2), absTolIn=(1e-30, 1e-30), maxItrIn=(1000, 1000)))


@pytest.fixture(scope="module")
def bind():
    if not OB.available():
        pytest.skip("oracle/_ref/libsvbind.so was not built (the reference sources were absent at build time)")
    _lib.load()
    return OB


def _dev(torch, a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def _ls(ls_type, prm):
    """The C-ABI solver settings equal to GM_PRM / NS_PRM."""
    from svfsiplus_amd.newton import ls_create
    if ls_type == _lib.LS_NS:
        ls = ls_create(_lib.LS_NS, relTol=1e-3, absTol=1e-30, maxItr=1000, dimKry=50, ns_gm=(1000, 1e-2),
                       ns_cg=(1000, 1e-2))
    else:
        ls = ls_create(ls_type, relTol=1e-6, absTol=1e-30, maxItr=1000, dimKry=50)
    got = (ls.RI.relTol, ls.RI.absTol, ls.RI.mItr, ls.RI.sD, ls.GM.relTol, ls.GM.absTol, ls.GM.mItr, ls.GM.sD,
           ls.CG.relTol, ls.CG.absTol, ls.CG.mItr)
    n = 11 if ls_type == _lib.LS_NS else 4  # GMRES reads only RI
    assert got[:n] == tuple(prm)[:n], (got, prm)
    return ls


def _capi(torch, p, ph, Ag, Yg, ls_type, prm, outlet=None, precision=0, coarse=0, system=None):
    """The same Newton iteration through the C ABI (svfsiplus_amd.newton)."""
    from svfsiplus_amd.newton import DeviceFace, DeviceLhs, DeviceMesh, b_assem_neu_bc, construct_fluid, fsils_solve
    L = DeviceLhs(p)
    objs = [L]
    if system is None:
        M = DeviceMesh(L, p)
        objs.append(M)
        dR = torch.zeros((p.nNo, 4), dtype=torch.float64, device="cuda")
        dV = torch.zeros((p.nnz, 16), dtype=torch.float64, device="cuda")
        dY = _dev(torch, Yg)
        construct_fluid(M, ph, _dev(torch, Ag), dY, dR, dV)
        if outlet is not None:
            Fa = DeviceFace(M, outlet[0], outlet[1])
            objs.append(Fa)
            b_assem_neu_bc(Fa, ph, dY, dR, dV, hg=None if outlet[2] is None else _dev(torch, outlet[2]))
    else:
        dR, dV = _dev(torch, system[0]), _dev(torch, system[1])
    ls = _ls(ls_type, prm)
    fsils_solve(L, ls, 4, dR, dV, precision=precision, coarse=coarse)
    x = dR.cpu().numpy()
    for o in reversed(objs):
        o.free()
    return x, ls


@pytest.mark.parametrize("ls_type", [_lib.LS_GMRES, _lib.LS_NS], ids=["gmres", "ns"])
def test_binding_against_reference_golden(torch_cuda, bind, ls_type):
    import _golden as G
    torch = torch_cuda
    p, ph, d = G.load("pipe_n4")
    prm = GM_PRM if ls_type == _lib.LS_GMRES else NS_PRM
    xg, tol = (d["gm_x"], float(d["gm_tol"])) if ls_type == _lib.LS_GMRES else (d["ns_x"], float(d["ns_tol"]))
    out = {}
    for mode in (0, 1):
        print(f"binding mode {mode} ls_type {ls_type}", flush=True)
        R, rec = bind.fluid_solve(p, ph, d["Ag0"], d["Yg0"], ls_type, prm, mode, iters=2)
        assert np.array_equal(R[0], R[1]) and rec[0] == rec[1]
        assert rec[0]["suc"] == 1.0 and rec[0]["device_assembly"] == float(mode)
        out[mode] = (R[0], rec[0])
    # solver only: the reference's own host assembly, solved on the device
    x0, ls0 = _capi(torch, p, ph, None, None, ls_type, prm, system=(d["R0"], d["Val0"]))
    r0 = out[0][1]
    print("solver-only binding itr", r0["itr"], "C ABI", ls0.RI.itr, "| rel to reference",
          np.linalg.norm(out[0][0] - xg) / np.linalg.norm(xg))
    assert np.array_equal(out[0][0], x0)
    assert r0["itr"] == ls0.RI.itr
    if ls_type == _lib.LS_NS:  # GMRES leaves GM / CG as fsils_ls_create left them (itr unset)
        assert (r0["GM_itr"], r0["CG_itr"], r0["Resm"], r0["Resc"]) == (ls0.GM.itr, ls0.CG.itr, ls0.Resm, ls0.Resc)
    assert r0["iNorm"] == ls0.RI.iNorm and r0["fNorm"] == ls0.RI.fNorm
    bar = 10 * tol if ls_type == _lib.LS_GMRES else 1e-3
    assert np.linalg.norm(out[0][0] - xg) <= bar * np.linalg.norm(xg)
    # device assembly: the C ABI's construct_fluid + fsils_solve on the same state
    x1, ls1 = _capi(torch, p, ph, d["Ag0"], d["Yg0"], ls_type, prm)
    r1 = out[1][1]
    print("device-assembly binding itr", r1["itr"], "C ABI", ls1.RI.itr, "| rel to solver-only",
          np.linalg.norm(out[1][0] - out[0][0]) / np.linalg.norm(out[0][0]))
    assert np.array_equal(out[1][0], x1) and r1["itr"] == ls1.RI.itr
    assert np.linalg.norm(out[1][0] - out[0][0]) <= bar * np.linalg.norm(out[0][0])


def test_binding_backflow_outlet_and_options(torch_cuda, bind):
    """The outlet's Neumann/backflow assembly through the back-end (velocity
    flowing back through part of the outlet, nonzero traction; golden
    backflow_n4's state) in both modes, FP64, then the mixed-precision
    two-level solve selected by the back-end's solve options."""
    import _golden as G
    torch = torch_cuda
    p, ph, d = G.load("backflow_n4")
    ph.beta, ph.qmTRI3 = float(d["beta"]), float(d["qmTRI3"])
    Ag = np.zeros((p.nNo, 4))
    Yg = d["Yg"]
    outlet = (d["faceIEN"], d["gE"], d["hg"])
    res = {}
    for mode in (0, 1):
        print(f"binding backflow mode {mode}", flush=True)
        R, rec = bind.fluid_solve(p, ph, Ag, Yg, _lib.LS_GMRES, GM_PRM, mode, iters=2, outlet=outlet)
        assert np.array_equal(R[0], R[1]) and rec[0]["suc"] == 1.0
        res[mode] = (R[0], rec[0])
    x1, ls1 = _capi(torch, p, ph, Ag, Yg, _lib.LS_GMRES, GM_PRM, outlet=outlet)
    rel = np.linalg.norm(res[1][0] - res[0][0]) / np.linalg.norm(res[0][0])
    print("backflow: itr host-assembly", res[0][1]["itr"], "device", res[1][1]["itr"], "C ABI", ls1.RI.itr,
          f"| device vs host assembly rel {rel:.2e}")
    assert np.array_equal(res[1][0], x1) and res[1][1]["itr"] == ls1.RI.itr
    assert rel <= 1e-5
    # options beyond the reference through the back-end
    print("binding backflow mode 1 mixed two-level", flush=True)
    R2, rec2 = bind.fluid_solve(p, ph, Ag, Yg, _lib.LS_GMRES, GM_PRM, 1, iters=1, outlet=outlet,
                                precision=_lib.PRECISION_MIXED, coarse=16)
    x2, ls2 = _capi(torch, p, ph, Ag, Yg, _lib.LS_GMRES, GM_PRM, outlet=outlet, precision=_lib.PRECISION_MIXED,
                    coarse=16)
    print("mixed two-level: itr", rec2[0]["itr"], "C ABI", ls2.RI.itr, "FP64", res[1][1]["itr"])
    assert np.array_equal(R2[0], x2) and rec2[0]["itr"] == ls2.RI.itr and rec2[0]["suc"] == 1.0
    # both stop on the 1e-6 residual; this state's system is ill-conditioned (observed 2.2e-5)
    assert np.linalg.norm(R2[0] - res[1][0]) <= 1e-4 * np.linalg.norm(res[1][0])
