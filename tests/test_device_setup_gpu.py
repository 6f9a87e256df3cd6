"""Mesh setup on the device (svg_lhs_create_dev, svg_msh_create_dev): the
block pattern equals the host build (lhsa.cpp:153-380) entry for entry; the
Jones-Plassmann colouring is a valid colouring (no two elements of a colour
share a node) with a stable colour order; the assembly through the device-
built slot map equals the host-built one to rounding (only the order of the
colour launches differs); a Newton time step on the device-built setup
matches the host-built one."""
import ctypes as C

import numpy as np
import pytest

from svfsiplus_amd import _lib, workload as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def phys():
    return W.Physics()


def _meshes():
    return [("pipe_n4", W.make_pipe(4)), ("pipe_n6_stenosed", W.make_pipe(6, sten_depth=0.4, sten_z0=1.25,
                                                                           sten_len=0.5)),
            ("bifurcation", W.make_bifurcation(4, 6, 6))]


@pytest.mark.parametrize("k", [0, 1, 2])
def test_device_pattern_and_colouring(torch_cuda, k):
    from svfsiplus_amd.newton import DeviceLhs, DeviceMesh, device_arrays
    name, p = _meshes()[k]
    lib = _lib.load()
    d_ien, d_x = device_arrays(p)
    L = DeviceLhs(p, d_ien=d_ien)
    assert L.nnz == p.nnz
    rp = np.empty(p.nNo + 1, dtype=np.int32)
    cp = np.empty(L.nnz, dtype=np.int32)
    _lib.check(lib.svg_lhs_pattern(L.h, rp.ctypes.data, cp.ctypes.data))
    assert np.array_equal(rp, p.rowPtr) and np.array_equal(cp, p.colPtr), name
    M = DeviceMesh(L, p, d_ien=d_ien, d_x=d_x)
    off = np.empty(M.nColors + 1, dtype=np.int32)
    _lib.check(lib.svg_msh_colors(M.h, off.ctypes.data))
    ien = np.empty((p.nEl, 4), dtype=np.int32)
    _lib.check(lib.svg_msh_sorted_ien(M.h, ien.ctypes.data))
    assert off[0] == 0 and off[-1] == p.nEl and np.all(np.diff(off) > 0)
    # a permutation of the elements, stable (ascending element ids) inside each colour
    key = {tuple(r): e for e, r in enumerate(np.asarray(p.IEN, dtype=np.int32))}
    ids = np.array([key[tuple(r)] for r in ien])
    assert np.array_equal(np.sort(ids), np.arange(p.nEl))
    for c in range(M.nColors):
        blk = ien[off[c]:off[c + 1]]
        nodes = blk.reshape(-1)
        assert np.unique(nodes).size == nodes.size, f"{name}: colour {c} has elements sharing a node"
        assert np.all(np.diff(ids[off[c]:off[c + 1]]) > 0)
    hc = C.c_int()
    col = np.empty(p.nEl, dtype=np.int32)
    _lib.check(lib.svg_color_elements(p.nNo, p.nEl, 4, np.ascontiguousarray(p.IEN, dtype=np.int32).ctypes.data,
                                      col.ctypes.data, C.byref(hc)))
    print(f"{name}: nEl {p.nEl} nnz {p.nnz}; colours: device {M.nColors}, host greedy {hc.value}")
    assert M.nColors <= 2 * hc.value


def test_device_setup_assembly_matches_host(torch_cuda, phys):
    from svfsiplus_amd.newton import DeviceLhs, DeviceMesh, construct_fluid, device_arrays
    import oracle
    torch = torch_cuda
    p = W.make_pipe(6)
    Ao, Yo, An, Yn = W.dirichlet_state(p, phys)
    Ag, Yg = oracle.pici(phys, Ao, An, Yo, Yn)
    dAg, dYg = (torch.as_tensor(a, device="cuda") for a in (Ag, Yg))
    out = []
    d_ien, d_x = device_arrays(p)
    for dev in (False, True):
        L = DeviceLhs(p, d_ien=d_ien if dev else None)
        M = DeviceMesh(L, p, d_ien=d_ien if dev else None, d_x=d_x if dev else None)
        R = torch.zeros((p.nNo, 4), dtype=torch.float64, device="cuda")
        V = torch.zeros((p.nnz, 16), dtype=torch.float64, device="cuda")
        construct_fluid(M, phys, dAg, dYg, R, V)
        torch.cuda.synchronize()
        out.append((R.cpu().numpy(), V.cpu().numpy()))
    (R0, V0), (R1, V1) = out
    # the assembly adds element contributions in element order whatever the
    # colouring, so both set-ups give the oracle's sums bit for bit
    Rr, Vr = oracle.construct_fluid(p, phys, Ag, Yg)
    assert np.array_equal(R0, Rr) and np.array_equal(V0, Vr)
    assert np.array_equal(R1, R0) and np.array_equal(V1, V0)


def test_device_setup_newton_step(torch_cuda, phys):
    from svfsiplus_amd.newton import FluidNewton, ls_create
    p = W.make_pipe(4)
    mk = lambda: ls_create(_lib.LS_GMRES, relTol=1e-6, absTol=1e-30, maxItr=1000, dimKry=50)
    h0 = FluidNewton(p, phys, mk()).time_step()
    fd = FluidNewton(p, phys, mk(), device_setup=True)
    h1 = fd.time_step()
    print("host setup  ", h0, "\ndevice setup", h1)
    # the first residual is the same (the assembly does not depend on the
    # set-up); the later ones follow linear solves stopped at a 1e-6 relative residual
    assert h0[0]["iNorm"] == h1[0]["iNorm"]
    for a, b in zip(h0, h1):
        assert abs(a["ls_itr"] - b["ls_itr"]) <= 2
        assert abs(a["iNorm"] - b["iNorm"]) <= 1e-6 * a["iNorm"]


def test_device_setup_rejects_bad_connectivity(torch_cuda):
    """A node id out of range is an error (checked on the device before any use),
    as the host pattern build reports it."""
    from svfsiplus_amd.newton import DeviceLhs
    p = W.make_pipe(2)
    ien = np.ascontiguousarray(p.IEN, dtype=np.int32).copy()
    ien[5, 2] = p.nNo  # one past the last node
    d_ien = torch_cuda.as_tensor(ien, device="cuda")
    with pytest.raises(RuntimeError, match="IEN entry out of range"):
        DeviceLhs(p, d_ien=d_ien)
    with pytest.raises(RuntimeError, match="IEN entry out of range"):
        W.pattern(p.nNo, ien)
