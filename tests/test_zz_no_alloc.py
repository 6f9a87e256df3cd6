"""The no-allocation dot product (svg_dot_v_noalloc) against svg_dot_v through
the C ABI: bit-identical results on two calls in a row, the second on a
smaller input (the static scratch keeps the first call's partial sums, which
the second must not pick up), and the size limit of its static scratch."""
import ctypes as C

import numpy as np
import pytest

from svfsiplus_amd import _lib

pytestmark = pytest.mark.gpu


def _dot(name, torch, dof, nNo, U, V):
    out = C.c_double(np.nan)
    s = torch.cuda.current_stream().cuda_stream
    _lib.call(name, dof, nNo, U.data_ptr(), V.data_ptr(), C.byref(out), s)
    return out.value


def test_noalloc_dot_equals_dot(torch_cuda):
    torch = torch_cuda
    g = torch.Generator(device="cuda").manual_seed(11)
    # 2^24 nodes x 4 dof = 2^26 entries (16384 partial sums), then 300 001 x 1
    # (293 partial sums): slots 293.. still hold the first call's partials
    for dof, nNo in ((4, 1 << 24), (1, 300001), (4, 0)):
        n = dof * nNo
        U = torch.randn(max(n, 1), generator=g, device="cuda", dtype=torch.float64)
        V = torch.randn(max(n, 1), generator=g, device="cuda", dtype=torch.float64)
        a = _dot("svg_dot_v_noalloc", torch, dof, nNo, U, V)
        b = _dot("svg_dot_v", torch, dof, nNo, U, V)
        print(f"dof {dof} nNo {nNo}: noalloc {a!r} dot {b!r}")
        assert np.float64(a).tobytes() == np.float64(b).tobytes()
        if n:
            exact = float(torch.dot(U.cpu(), V.cpu()))
            assert abs(a - exact) <= 1e-12 * float((U.abs() * V.abs()).sum())
        else:
            assert a == 0.0


def test_noalloc_dot_rejects_sizes_beyond_its_scratch(torch_cuda):
    torch = torch_cuda
    U = torch.zeros(4, device="cuda", dtype=torch.float64)
    out = C.c_double()
    s = torch.cuda.current_stream().cuda_stream
    # (1 << 26) + 1 nodes x 4 dof is one entry past 2^28: refused on the host
    # before any device work
    rc = _lib.load().svg_dot_v_noalloc(4, (1 << 26) + 1, U.data_ptr(), U.data_ptr(), C.byref(out), s)
    assert rc == _lib.SVG_ERR_ARG
    assert b"2^28" in _lib.load().svg_last_error()
