"""The drop-in back-end GpuFsilsLinearAlgebra (svfsiplus_amd/binding), driven
through the reference's own types and call sequence (test infrastructure only).

_ref/libsvbind.so is compiled by ref.mk from the binding's sources against the
reference's headers, with bind_harness.cpp and the reference's own objects
(fsils_lhs_create, fsils_bc_create, fsils_ls_create, ls_alloc, ls_solve,
construct_fluid, b_assem_neu_bc, ...), and linked against libsvgpu.so. The
harness runs the reference's Newton-iteration sequence with the back-end as
com_mod.eq[0].linear_algebra:

    mode 0   the reference's host construct_fluid / b_assem_neu_bc, whose
             element matrices reach the back-end through LinearAlgebra::assemble,
             then LinearAlgebra::solve on the device;
    mode 1   the back-end's device construct_fluid / b_assem_neu_bc (the branch
             a maintainer adds at eq_assem.cpp:400-402), then the solve.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
BIND_PATH = HERE / "_ref" / "libsvbind.so"

_bind = None


def available() -> bool:
    return BIND_PATH.exists()


def load():
    global _bind
    if _bind is None:
        lib = C.CDLL(str(BIND_PATH))
        lib.bind_last_error.restype = C.c_char_p
        lib.bind_fluid_solve.restype = C.c_int
        lib.bind_fluid_solve.argtypes = ([C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                          C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p])
        _bind = lib
    return _bind


def _p(a):
    return None if a is None else a.ctypes.data


def fluid_solve(pipe, phys, Ag, Yg, ls_type: int, ls_prm, mode: int, iters: int = 1, precision: int = 0,
                coarse: int = 0, outlet=None):
    """Runs `iters` Newton iterations of the same state through the back-end.
    outlet = (faceIEN (nFe, 3), gE (nFe,), hg (nNo,) or None): a Neumann face
    assembled with backflow (phys.beta, phys.qmTRI3). The lhs faces are
    pipe.faces, Dirichlet. Returns (R (iters, nNo, 4), records)."""
    lib = load()
    globs = [np.ascontiguousarray(f[1], dtype=np.int32) for f in pipe.faces]
    nn = np.array([len(g) for g in globs], dtype=np.int32)
    glob = np.ascontiguousarray(np.concatenate(globs), dtype=np.int32)
    val = np.zeros((glob.shape[0], 3))
    prm = np.array([phys.rho, phys.mu, phys.f[0], phys.f[1], phys.f[2], phys.K_darcy, phys.af, phys.am, phys.gam,
                    phys.dt])
    x = np.ascontiguousarray(pipe.x, dtype=np.float64)
    IEN = np.ascontiguousarray(pipe.IEN, dtype=np.int32)
    Ag = np.ascontiguousarray(Ag, dtype=np.float64)
    Yg = np.ascontiguousarray(Yg, dtype=np.float64)
    fi = ge = hg = None
    nFe = 0
    if outlet is not None:
        fi = np.ascontiguousarray(outlet[0], dtype=np.int32)
        ge = np.ascontiguousarray(outlet[1], dtype=np.int32)
        hg = None if outlet[2] is None else np.ascontiguousarray(outlet[2], dtype=np.float64)
        nFe = fi.shape[0]
    lp = np.array(ls_prm, dtype=np.float64)
    iopt = np.array([mode, iters, precision, coarse], dtype=np.int32)
    R = np.zeros((iters, pipe.nNo, 4))
    out = np.zeros((iters, 10))
    rc = lib.bind_fluid_solve(pipe.nEl, _p(IEN), pipe.nNo, _p(x), _p(Ag), _p(Yg), _p(prm), pipe.nnz, _p(pipe.rowPtr),
                              _p(pipe.colPtr), len(globs), _p(nn), _p(glob), _p(val), nFe, _p(fi), _p(ge), _p(hg),
                              float(phys.beta), float(phys.qmTRI3), ls_type, _p(lp), _p(iopt), _p(R), _p(out))
    if rc:
        raise RuntimeError(lib.bind_last_error().decode())
    keys = ["itr", "iNorm", "fNorm", "suc", "dB", "GM_itr", "CG_itr", "Resm", "Resc", "device_assembly"]
    return R, [dict(zip(keys, o.tolist())) for o in out]
