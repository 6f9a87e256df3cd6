"""Synthetic pipe workloads of the fluid Newton step (host-side setup).

A straight pipe (radius 0.3, length 3, axis z) meshed by the native generator
(svg_pipe_generate): n x n hex cross-section mapped to the disc, nz = 5n hexes
axially, 6 tets per hex, interior nodes jittered by U(-0.1 h, 0.1 h) from a
seeded splitmix64 stream. Boundary conditions follow the reference's pipe
cases: parabolic Dirichlet inflow at z = 0, no-slip Dirichlet walls, a
zero-traction outlet (natural; no FSILS face, baf_ini.cpp:752-756).

Sizes: n=8 -> 15 360 tets / 3 321 nodes; n=34 -> 1 179 120 tets / 209 475 nodes.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib

SEED_MESH = 20261202


@dataclass
class Physics:
    rho: float = 1.06
    mu: float = 0.04
    peak: float = 30.0          # inflow centre-line velocity
    dt: float = 1.0e-3
    ro_inf: float = 0.5         # Spectral_radius_of_infinite_time_step
    f: tuple = (0.0, 0.0, 0.0)
    K_darcy: float = 0.0
    qmTET4: float = (5.0 + 3.0 * math.sqrt(5.0)) / 20.0
    beta: float = 0.2           # Backflow_stabilization_coefficient (Parameters.cpp:1390 default)
    qmTRI3: float = 2.0 / 3.0   # TRI3 face quadrature modifier (ComMod.h:611)

    # generalized-alpha for a first-order equation (initialize.cpp:430-463)
    @property
    def am(self) -> float:
        return 0.5 * (3.0 - self.ro_inf) / (1.0 + self.ro_inf)

    @property
    def af(self) -> float:
        return 1.0 / (1.0 + self.ro_inf)

    @property
    def gam(self) -> float:
        return 0.5 + self.am - self.af

    def fluid_params(self) -> _lib.FluidParams:
        p = _lib.FluidParams()
        p.rho, p.mu = self.rho, self.mu
        p.f[0], p.f[1], p.f[2] = self.f
        p.K_darcy = self.K_darcy
        p.af, p.am, p.gam, p.dt = self.af, self.am, self.gam, self.dt
        p.qmTET4 = self.qmTET4
        return p


@dataclass
class Pipe:
    n: int
    nz: int
    x: np.ndarray          # (nNo, 3) row-major == reference x(3,nNo) column-major
    IEN: np.ndarray        # (nEl, 4)
    flags: np.ndarray      # (nNo,)
    rowPtr: np.ndarray     # (nNo+1,)
    colPtr: np.ndarray     # (nnz,)
    radius: float
    length: float
    faces: list = field(default_factory=list)   # [(name, node ids)] Dirichlet faces in FSILS order

    @property
    def nNo(self) -> int:
        return self.x.shape[0]

    @property
    def nEl(self) -> int:
        return self.IEN.shape[0]

    @property
    def nnz(self) -> int:
        return self.colPtr.shape[0]

    def rowPtr2(self) -> np.ndarray:
        """FSILS inclusive [s, e] pairs, (nNo, 2) == reference rowPtr(2,nNo)."""
        r = np.empty((self.nNo, 2), dtype=np.int32)
        r[:, 0] = self.rowPtr[:-1]
        r[:, 1] = self.rowPtr[1:] - 1
        return r

    def diagPtr(self) -> np.ndarray:
        rows = np.repeat(np.arange(self.nNo, dtype=np.int64), np.diff(self.rowPtr))
        d = np.empty(self.nNo, dtype=np.int32)
        hit = np.nonzero(self.colPtr == rows)[0]
        d[rows[hit]] = hit
        return d


def make_pipe(n: int, nz: int | None = None, radius: float = 0.3, length: float = 3.0, jitter: float = 0.1,
              seed: int = SEED_MESH, sten_depth: float = 0.0, sten_z0: float = 0.0, sten_len: float = 0.0) -> Pipe:
    lib = _lib.load()
    nz = 5 * n if nz is None else nz
    sp = _lib.PipeSpec(n, nz, radius, length, jitter, seed, sten_depth, sten_z0, sten_len)
    nNo, nEl = C.c_int(), C.c_int()
    _lib.check(lib.svg_pipe_counts(C.byref(sp), C.byref(nNo), C.byref(nEl)))
    x = np.empty((nNo.value, 3), dtype=np.float64)
    IEN = np.empty((nEl.value, 4), dtype=np.int32)
    flags = np.empty(nNo.value, dtype=np.int32)
    _lib.check(lib.svg_pipe_generate(C.byref(sp), x.ctypes.data, IEN.ctypes.data, flags.ctypes.data))
    nnz = C.c_int()
    _lib.check(lib.svg_csr_pattern_count(nNo.value, nEl.value, 4, IEN.ctypes.data, C.byref(nnz)))
    rowPtr = np.empty(nNo.value + 1, dtype=np.int32)
    colPtr = np.empty(nnz.value, dtype=np.int32)
    _lib.check(lib.svg_csr_pattern(nNo.value, nEl.value, 4, IEN.ctypes.data, rowPtr.ctypes.data,
                                   colPtr.ctypes.data))
    inlet = np.nonzero(flags & _lib.NODE_INLET)[0].astype(np.int32)
    wall = np.nonzero(flags & _lib.NODE_WALL)[0].astype(np.int32)
    p = Pipe(n, nz, x, IEN, flags, rowPtr, colPtr, radius, length)
    p.faces = [("inlet", inlet), ("wall", wall)]
    return p


def pattern(nNo: int, IEN: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Block-CSR pattern of a TET4 mesh (what lhsa builds, lhsa.cpp:153-380):
    rowPtr (nNo+1, exclusive end) and the sorted colPtr."""
    lib = _lib.load()
    IEN = np.ascontiguousarray(IEN, dtype=np.int32)
    nEl = IEN.shape[0]
    nnz = C.c_int()
    _lib.check(lib.svg_csr_pattern_count(nNo, nEl, 4, IEN.ctypes.data, C.byref(nnz)))
    rowPtr = np.empty(nNo + 1, dtype=np.int32)
    colPtr = np.empty(nnz.value, dtype=np.int32)
    _lib.check(lib.svg_csr_pattern(nNo, nEl, 4, IEN.ctypes.data, rowPtr.ctypes.data, colPtr.ctypes.data))
    return rowPtr, colPtr


def _finish_mesh(x, IEN, flags, n, nz, radius, length) -> Pipe:
    rowPtr, colPtr = pattern(x.shape[0], IEN)
    inlet = np.nonzero(flags & _lib.NODE_INLET)[0].astype(np.int32)
    wall = np.nonzero(flags & _lib.NODE_WALL)[0].astype(np.int32)
    p = Pipe(n, nz, x, IEN, flags, rowPtr, colPtr, radius, length)
    p.faces = [("inlet", inlet), ("wall", wall)]
    return p


def make_bifurcation(n: int, nzp: int, nzd: int, radius: float = 0.3, length_parent: float = 1.5,
                     length_daughter: float = 1.5, angle0: float = 0.6, angle1: float = 0.6, split: float = 0.5,
                     jitter: float = 0.1, seed: int = SEED_MESH) -> Pipe:
    """A bifurcating pipe (svg_bif_generate): parent + two Murray-law daughters."""
    lib = _lib.load()
    sp = _lib.BifSpec(n, nzp, nzd, radius, length_parent, length_daughter, angle0, angle1, split, jitter, seed)
    nNo, nEl = C.c_int(), C.c_int()
    _lib.check(lib.svg_bif_counts(C.byref(sp), C.byref(nNo), C.byref(nEl)))
    x = np.empty((nNo.value, 3), dtype=np.float64)
    IEN = np.empty((nEl.value, 4), dtype=np.int32)
    flags = np.empty(nNo.value, dtype=np.int32)
    _lib.check(lib.svg_bif_generate(C.byref(sp), x.ctypes.data, IEN.ctypes.data, flags.ctypes.data))
    return _finish_mesh(x, IEN, flags, n, nzp + nzd, radius, length_parent)


def bifurcation_batch(count: int = 256, seed: int = 20261202, tets_min: float = 5e3, tets_max: float = 8e4,
                      radius: float = 0.3):
    """Seeded specs of the batched workload: tet counts log-uniform in
    [tets_min, tets_max] (cross-section n = 4, 8 or 12 by size, layers from the
    count, segment lengths keeping the axial spacing at 1.5 cross-section cells),
    daughter angles uniform in [25, 45] degrees, Murray split uniform in
    [0.35, 0.65]. Returns keyword dicts for make_bifurcation."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        T = math.exp(rng.uniform(math.log(tets_min), math.log(tets_max)))
        n = 4 if T < 1.2e4 else (8 if T < 5e4 else 12)
        layers = max(6, int(round(T / (6.0 * n * n))))   # parent + daughter layers, n^2 hexes each
        nzp = max(2, layers // 2)
        nzd = max(3, layers - nzp)
        hz = 1.5 * 2.0 * radius / n
        a0, a1 = np.radians(rng.uniform(25.0, 45.0, 2))
        out.append(dict(n=n, nzp=nzp, nzd=nzd, radius=radius, length_parent=nzp * hz, length_daughter=nzd * hz,
                        angle0=float(a0), angle1=float(a1), split=float(rng.uniform(0.35, 0.65)),
                        seed=int(rng.integers(1, 2**31))))
    return out


def dirichlet_state(p: Pipe, phys: Physics):
    """Initial (Ao, Yo, An, Yn) of the first time step: Ao = Yo = 0, then the
    predictor (picp, pic.cpp:591-715: An = Ao (gam-1)/gam, Yn = Yo) and the
    strong Dirichlet values (set_bc_dir_l, set_bc.cpp:1045-1053, steady:
    A = 0, Y = g gx n), applied face by face in FSILS face order: the inlet
    gets u_z = peak (1 - r^2/R^2), the wall (which contains the inlet and
    outlet rings) gets u = 0."""
    nNo = p.nNo
    Ao = np.zeros((nNo, 4))
    Yo = np.zeros((nNo, 4))
    coef = (phys.gam - 1.0) / phys.gam
    An = Ao * coef
    Yn = Yo.copy()
    inlet = p.faces[0][1]
    r2 = (p.x[inlet, 0] ** 2 + p.x[inlet, 1] ** 2) / (p.radius ** 2)
    An[inlet, 0:3] = 0.0
    Yn[inlet, 0:3] = 0.0
    Yn[inlet, 2] = phys.peak * (1.0 - r2)
    wall = p.faces[1][1]
    An[wall, 0:3] = 0.0
    Yn[wall, 0:3] = 0.0
    return Ao, Yo, An, Yn


def dirichlet_faces(p: Pipe, phys: Physics):
    """The strong Dirichlet faces in FSILS face order, as set_bc_dir takes them:
    [(gN, gxn(nNo, 3))] with dirA = 0, dirY = 1 (steady values): the inlet's
    parabola u_z = peak (1 - r^2/R^2), then the wall (u = 0)."""
    inlet = np.ascontiguousarray(p.faces[0][1], dtype=np.int32)
    r2 = (p.x[inlet, 0] ** 2 + p.x[inlet, 1] ** 2) / (p.radius ** 2)
    g_in = np.zeros((inlet.shape[0], 3))
    g_in[:, 2] = phys.peak * (1.0 - r2)
    wall = np.ascontiguousarray(p.faces[1][1], dtype=np.int32)
    return [(inlet, g_in), (wall, np.zeros((wall.shape[0], 3)))]


def face_arrays(p: Pipe):
    """Concatenated FSILS face data: (nNo per face, glob, val(3,n) with val = 0)."""
    nn = np.array([len(f[1]) for f in p.faces], dtype=np.int32)
    glob = np.concatenate([f[1] for f in p.faces]).astype(np.int32)
    val = np.zeros((glob.shape[0], 3))
    return nn, glob, val


def face_normal_integrals(p: Pipe, faceIEN, inner):
    """The FSILS face data of a coupled Neumann face: its nodes and
    val(a) = integral of N_a n over the face (n outward, N_a the TRI3 shape
    function: area n / 3 per triangle and vertex), the vector a resistance
    boundary condition couples through (fsils_bc_create of the face with
    BC_TYPE_Neu). Returns (nodes, val (nNo, 3))."""
    nodes, inv = np.unique(np.asarray(faceIEN).ravel(), return_inverse=True)
    x = p.x
    a, b, c = (x[faceIEN[:, k]] for k in range(3))
    cr = 0.5 * np.cross(b - a, c - a)                       # area-weighted normal
    out = np.sign(np.einsum("ij,ij->i", cr, a - x[inner]))  # away from the interior node
    cr = cr * out[:, None]
    val = np.zeros((nodes.shape[0], 3))
    np.add.at(val, inv.reshape(-1, 3)[:, 0], cr / 3.0)
    np.add.at(val, inv.reshape(-1, 3)[:, 1], cr / 3.0)
    np.add.at(val, inv.reshape(-1, 3)[:, 2], cr / 3.0)
    return nodes.astype(np.int32), val


def outlet_faces(p: Pipe, flag: int = _lib.NODE_OUTLET):
    """TRI3 faces of the outlet (every tet face whose three nodes carry the
    outlet flag, or `flag`): faceIEN (nFe, 3) in the tet's node order, the
    parent tet gE (nFe,) and its fourth node (nFe,), which orients the outward
    normal as gnnb does."""
    out = (p.flags & flag) != 0
    faces, parent, inner = [], [], []
    for k in range(4):
        keep = [a for a in range(4) if a != k]
        tri = p.IEN[:, keep]
        hit = np.all(out[tri], axis=1)
        faces.append(tri[hit])
        parent.append(np.nonzero(hit)[0])
        inner.append(p.IEN[hit, k])
    return (np.ascontiguousarray(np.concatenate(faces), dtype=np.int32),
            np.ascontiguousarray(np.concatenate(parent), dtype=np.int32),
            np.ascontiguousarray(np.concatenate(inner), dtype=np.int32))
