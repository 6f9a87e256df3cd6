"""svfsiplus_amd — MI355X-native fluid Newton step for svFSIplus.

The hot path (TET4 P1-P1 VMS assembly + FSILS Krylov solve) runs in
libsvgpu.so, hand-written HIP for gfx950 behind the C ABI of
include/svfsi_gpu.h. This package holds the ctypes binding (_lib), the
synthetic pipe workloads (workload) and the host driver that mirrors the
reference's Newton loop (newton).
"""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
