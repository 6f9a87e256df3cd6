"""CPU tests of the host side: the C-ABI library loads and exports every
declared entry point; the pipe generator, CSR pattern and element colouring."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from svfsiplus_amd import _lib, workload as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "svfsi_gpu.h")).read()
    return sorted(set(re.findall(r"\b(svg_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.EXPORTED)


def test_pipe_sizes():
    p = W.make_pipe(8)
    assert (p.nEl, p.nNo) == (15360, 3321)
    sp = _lib.PipeSpec(34, 170, 0.3, 3.0, 0.1, W.SEED_MESH, 0, 0, 0)
    nNo, nEl = C.c_int(), C.c_int()
    _lib.check(_lib.load().svg_pipe_counts(C.byref(sp), C.byref(nNo), C.byref(nEl)))
    assert (nEl.value, nNo.value) == (1179120, 209475)
    with pytest.raises(_lib.SvgError):
        _lib.check(_lib.load().svg_pipe_counts(C.byref(_lib.PipeSpec(3, 5, 0.3, 3, 0, 1, 0, 0, 0)),
                                               C.byref(nNo), C.byref(nEl)))


def test_pipe_geometry_and_orientation():
    p = W.make_pipe(6)
    x, IEN = p.x, p.IEN
    v = [x[IEN[:, a]] for a in range(4)]
    det = np.einsum("ij,ij->i", v[1] - v[0], np.cross(v[2] - v[0], v[3] - v[0]))
    assert (det > 0).all()
    # the tets tile the pipe: total volume close to pi R^2 L (polygonal cross-section)
    vol = det.sum() / 6.0
    assert abs(vol - np.pi * 0.3 ** 2 * 3.0) < 0.02 * vol
    r = np.hypot(x[:, 0], x[:, 1])
    wall = (p.flags & _lib.NODE_WALL) != 0
    assert np.allclose(r[wall], 0.3, rtol=1e-12)
    assert (r[~wall] < 0.3).all()
    assert np.allclose(x[(p.flags & _lib.NODE_INLET) != 0, 2], 0.0)
    assert np.allclose(x[(p.flags & _lib.NODE_OUTLET) != 0, 2], 3.0)
    # seeded: same seed -> same mesh, other seed -> different interior
    q = W.make_pipe(6)
    assert np.array_equal(p.x, q.x)
    assert not np.array_equal(p.x, W.make_pipe(6, seed=1).x)


def test_pattern_matches_element_pairs():
    p = W.make_pipe(4, 6)
    pairs = set()
    for e in p.IEN:
        for a in e:
            for b in e:
                pairs.add((int(a), int(b)))
    rows = np.repeat(np.arange(p.nNo), np.diff(p.rowPtr))
    got = set(zip(rows.tolist(), p.colPtr.tolist()))
    assert got == pairs
    for a in range(p.nNo):
        cols = p.colPtr[p.rowPtr[a]:p.rowPtr[a + 1]]
        assert (np.diff(cols) > 0).all()
    assert (p.colPtr[p.diagPtr()] == np.arange(p.nNo)).all()


def test_element_colouring_is_valid():
    p = W.make_pipe(6)
    col = np.empty(p.nEl, dtype=np.int32)
    nc = C.c_int()
    _lib.check(_lib.load().svg_color_elements(p.nNo, p.nEl, 4, p.IEN.ctypes.data, col.ctypes.data, C.byref(nc)))
    assert 24 <= nc.value <= 64
    for c in range(nc.value):
        nodes = p.IEN[col == c].reshape(-1)
        assert len(nodes) == len(np.unique(nodes)), f"colour {c} shares a node"


def test_stenosis_narrows_the_pipe():
    p = W.make_pipe(4, 20, sten_depth=0.5, sten_z0=1.0, sten_len=0.5)
    wall = (p.flags & _lib.NODE_WALL) != 0
    r = np.hypot(p.x[wall, 0], p.x[wall, 1])
    z = p.x[wall, 2]
    mid = np.abs(z - 1.25) < 1e-9
    assert np.allclose(r[mid], 0.15, rtol=1e-12)
    assert np.allclose(r[z < 0.99], 0.3, rtol=1e-12)


def test_cpp_driver_built_and_usage():
    """The C++ Newton-step driver links against the library and checks its
    arguments without touching the GPU."""
    import subprocess
    exe = os.path.join(os.path.dirname(_lib.LIB_PATH), "svg_newton_step")
    assert os.path.exists(exe), "run python -m svfsiplus_amd.build"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def test_bifurcation_mesh_valid_and_union():
    """Bifurcation generator: positive tets, a manifold conforming mesh whose
    boundary faces are exactly inlet, wall and the two outlets; the batch union
    keeps each mesh's pattern block-diagonal."""
    from collections import Counter
    from svfsiplus_amd import batch as B
    for sp in W.bifurcation_batch(3, seed=3, tets_max=2e4):
        p = W.make_bifurcation(**sp)
        x, T = p.x, p.IEN
        det = np.einsum("ij,ij->i", x[T[:, 1]] - x[T[:, 0]],
                        np.cross(x[T[:, 2]] - x[T[:, 0]], x[T[:, 3]] - x[T[:, 0]]))
        assert det.min() > 0
        faces = Counter()
        for k in range(4):
            for f in map(tuple, np.sort(np.delete(T, k, axis=1), axis=1)):
                faces[f] += 1
        assert max(faces.values()) == 2
        kinds = Counter()
        for f, c in faces.items():
            if c == 1:
                fl = p.flags[list(f)]
                kinds["inlet" if np.all(fl & 1) else "outlet" if np.all(fl & 4) else
                      "wall" if np.all(fl & 2) else "other"] += 1
        assert kinds["other"] == 0 and kinds["inlet"] == 2 * sp["n"] ** 2 and kinds["outlet"] == 2 * sp["n"] ** 2
    ms = B.make_batch(W.bifurcation_batch(3, seed=4, tets_max=1e4))
    u, off = B.union(ms)
    nnz = C.c_int()
    _lib.check(_lib.load().svg_csr_pattern_count(u.nNo, u.nEl, 4, u.IEN.ctypes.data, C.byref(nnz)))
    rp = np.empty(u.nNo + 1, np.int32)
    cp = np.empty(nnz.value, np.int32)
    _lib.check(_lib.load().svg_csr_pattern(u.nNo, u.nEl, 4, u.IEN.ctypes.data, rp.ctypes.data, cp.ctypes.data))
    assert np.array_equal(rp, u.rowPtr) and np.array_equal(cp, u.colPtr)
    assert off[-1] == u.nNo


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def _ls_fields(ls):
    return {name: {f: getattr(getattr(ls, name), f) for f in ("relTol", "absTol", "mItr", "sD")}
            for name in ("RI", "GM", "CG")}


def test_ls_create_defaults_match_reference():
    """svg_ls_create (host code, no device call) against the reference's own
    fsils_ls_create (ls.cpp:40-120) for every LS type: the fixture
    tests/golden/ls_defaults.json written from it, and the reference compiled
    here when present. Fields the reference leaves unset (-7777 in the fixture)
    are 0 here."""
    import json
    import oracle
    with open(os.path.join(ROOT, "tests", "golden", "ls_defaults.json")) as f:
        golden = json.load(f)
    live = {t: oracle.ref_ls_create(int(t)) for t in golden} if oracle.ref_available() else {}
    assert set(golden) == {str(t) for t in (_lib.LS_NS, _lib.LS_GMRES, _lib.LS_CG, _lib.LS_BICGS)}
    for t, ref in golden.items():
        if t in live:
            assert live[t] == ref, t
        ls = _lib.Ls()
        _lib.call("svg_ls_create", C.byref(ls), int(t))
        assert ls.LS_type == ref["LS_type"] == int(t)
        got = _ls_fields(ls)
        for name in ("RI", "GM", "CG"):
            for f, r in ref[name].items():
                assert got[name][f] == (0 if r == -7777.0 else r), (t, name, f)
