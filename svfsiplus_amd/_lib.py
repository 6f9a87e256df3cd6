"""ctypes binding of the native library libsvgpu.so (the C ABI in include/svfsi_gpu.h).

The library is loaded from the package directory only; there is no fallback:
if it is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libsvgpu.so"

SVG_OK = 0
SVG_ERR_ARG = 1
SVG_ERR_HIP = 2
SVG_ERR_JACOBIAN = 3
SVG_ERR_SINGULAR = 4
SVG_ERR_ILLCOND = 5
SVG_ERR_LSTYPE = 6
SVG_ERR_NEU_RES = 7
SVG_ERR_UNSUPPORTED = 8

LS_BICGS, LS_NS, LS_GMRES, LS_CG = 795, 796, 797, 798
PREC_NONE, PREC_FSILS = 700, 701
BC_DIR, BC_NEU = 0, 1
NODE_INLET, NODE_WALL, NODE_OUTLET = 1, 2, 4


class SvgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[svg error {code}] {msg}")
        self.code = code


class PipeSpec(C.Structure):
    _fields_ = [("n", C.c_int), ("nz", C.c_int), ("radius", C.c_double), ("length", C.c_double),
                ("jitter", C.c_double), ("seed", C.c_ulonglong), ("sten_depth", C.c_double),
                ("sten_z0", C.c_double), ("sten_len", C.c_double)]


class BifSpec(C.Structure):
    _fields_ = [("n", C.c_int), ("nzp", C.c_int), ("nzd", C.c_int), ("radius", C.c_double),
                ("length_parent", C.c_double), ("length_daughter", C.c_double), ("angle0", C.c_double),
                ("angle1", C.c_double), ("split", C.c_double), ("jitter", C.c_double), ("seed", C.c_ulonglong)]


class FluidParams(C.Structure):
    _fields_ = [("rho", C.c_double), ("mu", C.c_double), ("f", C.c_double * 3), ("K_darcy", C.c_double),
                ("af", C.c_double), ("am", C.c_double), ("gam", C.c_double), ("dt", C.c_double),
                ("qmTET4", C.c_double)]


class SubLs(C.Structure):
    _fields_ = [("suc", C.c_int), ("mItr", C.c_int), ("sD", C.c_int), ("itr", C.c_int), ("cM", C.c_int),
                ("cN", C.c_int), ("cD", C.c_int), ("absTol", C.c_double), ("relTol", C.c_double),
                ("iNorm", C.c_double), ("fNorm", C.c_double), ("dB", C.c_double), ("callD", C.c_double)]


class SolveOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("method", C.c_int), ("fallback_window", C.c_int),
                ("fallback_itr", C.c_int), ("net_period", C.c_int), ("coarse", C.c_int), ("coarse_keep", C.c_int), ("reserved", C.c_int),
                ("lpc", C.c_void_p)]


PRECISION_FP64, PRECISION_MIXED = 0, 1
COARSE_APPLY, COARSE_RESIDUAL, COARSE_SMOOTH = 0, 1, 2
METHOD_PLAIN, METHOD_LEARNED = 0, 1


class Ls(C.Structure):
    _fields_ = [("LS_type", C.c_int), ("Resm", C.c_int), ("Resc", C.c_int), ("GM", SubLs), ("CG", SubLs),
                ("RI", SubLs)]


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)

_SIGS = {
    "svg_last_error": (C.c_char_p, []),
    "svg_pipe_counts": (_I, [C.POINTER(PipeSpec), _IP, _IP]),
    "svg_pipe_generate": (_I, [C.POINTER(PipeSpec), _P, _P, _P]),
    "svg_bif_counts": (_I, [C.POINTER(BifSpec), _IP, _IP]),
    "svg_bif_generate": (_I, [C.POINTER(BifSpec), _P, _P, _P]),
    "svg_csr_pattern_count": (_I, [_I, _I, _I, _P, _IP]),
    "svg_csr_pattern": (_I, [_I, _I, _I, _P, _P, _P]),
    "svg_color_elements": (_I, [_I, _I, _I, _P, _P, _IP]),
    "svg_lhs_create": (_I, [_I, _I, _P, _P, _I, C.POINTER(_P)]),
    "svg_lhs_bc_create": (_I, [_P, _I, _I, _I, _I, _P, _P]),
    "svg_lhs_free": (_I, [_P]),
    "svg_msh_create": (_I, [_P, _I, _P, _I, _P, C.POINTER(_P)]),
    "svg_msh_free": (_I, [_P]),
    "svg_msh_num_colors": (_I, [_P, _IP]),
    "svg_lhs_create_dev": (_I, [_I, _I, _P, _I, C.POINTER(_P), _P]),
    "svg_lhs_nnz": (_I, [_P, _IP]),
    "svg_lhs_pattern": (_I, [_P, _P, _P]),
    "svg_msh_create_dev": (_I, [_P, _I, _P, _P, C.POINTER(_P), _P]),
    "svg_msh_colors": (_I, [_P, _P]),
    "svg_msh_sorted_ien": (_I, [_P, _P]),
    "svg_construct_fluid": (_I, [_P, C.POINTER(FluidParams), _P, _P, _P, _P, _P, _P]),
    "svg_face_create": (_I, [_P, _I, _P, _P, C.POINTER(_P)]),
    "svg_face_free": (_I, [_P]),
    "svg_b_assem_neu_bc": (_I, [_P, C.POINTER(FluidParams), _D, _D, _P, _P, _P, _P, _P]),
    "svg_ls_create": (_I, [C.POINTER(Ls), _I]),
    "svg_fsils_solve": (_I, [_P, C.POINTER(Ls), _I, _P, _P, _I, _I, _P, _P, _P]),
    "svg_fsils_solve_ex": (_I, [_P, C.POINTER(Ls), _I, _P, _P, _I, _I, _P, _P, C.POINTER(SolveOpts), _P]),
    "svg_spar_mul_vv": (_I, [_P, _I, _P, _P, _P, _P]),
    "svg_precond_diag": (_I, [_P, _I, _P, _P, _P, _P]),
    "svg_spmv32": (_I, [_P, _P, _P, _P, _P]),
    "svg_spmv32_tune": (_I, [_P, _P, _P, _P, _IP, _P]),
    "svg_spmv_variants": (_I, [_P, _IP, _IP]),
    "svg_spmv_num_variants": (_I, [_IP]),
    "svg_spmv_time_variant": (_I, [_P, _I, _P, _P, _P, _I, _DP, _P]),
    "svg_spmv_tune": (_I, [_P, _P, _IP, _P]),
    "svg_dot_v": (_I, [_I, _I, _P, _P, _DP, _P]),
    "svg_norm_v": (_I, [_I, _I, _P, _DP, _P]),
    "svg_dot_v_noalloc": (_I, [_I, _I, _P, _P, _DP, _P]),
    "svg_coarse_prepare": (_I, [_P, _I, _P, _P]),
    "svg_coarse_apply": (_I, [_P, _I, _I, _P, _P, _P]),
    "svg_coarse_info": (_I, [_P, _IP, _P]),
    "svg_pici": (_I, [_I, _D, _D, _P, _P, _P, _P, _P, _P, _P]),
    "svg_picc_update": (_I, [_I, _D, _D, _P, _P, _P, _P]),
    "svg_fill": (_I, [_P, C.c_size_t, _D, _P]),
    "svg_picp": (_I, [_I, _D, _P, _P, _P, _P, _P]),
    "svg_set_bc_dir": (_I, [_I, _P, _I, _I, _I, C.c_uint, _P, _D, _D, _P, _P, _P]),
    "svg_prof_enable": (_I, [_I]),
    "svg_prof_only": (_I, [C.c_char_p]),
    "svg_prof_sample": (_I, [_I]),
    "svg_prof_reset": (_I, []),
    "svg_prof_get": (_I, [C.c_char_p, _IP, _DP]),
    "svg_prof_get_bytes": (_I, [C.c_char_p, _DP]),
    "svg_batch_create": (_I, [_P, _I, _P, _I, C.POINTER(_P)]),
    "svg_batch_free": (_I, [_P]),
    "svg_batch_gmres": (_I, [_P, C.POINTER(SubLs), _P, _P, _I, C.POINTER(SubLs), _P]),
    "svg_batch_fsils_solve": (_I, [_P, C.POINTER(SubLs), _P, _P, _I, C.POINTER(SubLs), _P]),
    "svg_batch_newton_iteration": (_I, [_P, _P, _P, C.POINTER(FluidParams), _D, _D, _P, _P, _P, _P, _P, _P, _P, _P,
                                        C.POINTER(SubLs), C.POINTER(SolveOpts), C.POINTER(SubLs), _P]),
    "svg_batch_gmres_ex": (_I, [_P, C.POINTER(SubLs), _P, _P, C.POINTER(SolveOpts), C.POINTER(SubLs), _P]),
    "svg_batch_fsils_solve_ex": (_I, [_P, C.POINTER(SubLs), _P, _P, C.POINTER(SolveOpts), C.POINTER(SubLs), _P]),
    "svg_batch_fallback_itr": (_I, [_P, _P]),
    "svg_batch_replays": (_I, [_P, _P]),
    "svg_lpc_num_params": (_I, [_IP]),
    "svg_lpc_create": (_I, [C.POINTER(_P)]),
    "svg_lpc_free": (_I, [_P]),
    "svg_lpc_set_params": (_I, [_P, _P]),
    "svg_lpc_get_params": (_I, [_P, _P]),
    "svg_lpc_prepare": (_I, [_P, _P, _P]),
    "svg_lpc_apply": (_I, [_P, _P, _I, _P, _P, _P]),
    "svg_lpc_loss_grad": (_I, [_P, _P, _P, _I, _P, _DP, _P, _P]),
    "svg_lpc_train_step": (_I, [_P, _P, _P, _I, _P, C.c_float, _DP, _P]),
    "svg_lpc_train_step_ex": (_I, [_P, _P, _P, _I, _P, _P, C.c_float, _DP, _P]),
}

EXPORTED = tuple(_SIGS)

_lib = None


def load() -> C.CDLL:
    """Loads libsvgpu.so from the package directory (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m svfsiplus_amd.build`")
    # Device memory and streams come from torch: load its HIP runtime first so
    # that the library binds to the same runtime instance in this process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != SVG_OK:
        msg = load().svg_last_error().decode(errors="replace")
        raise SvgError(rc, msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))
