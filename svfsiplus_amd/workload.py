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


def face_arrays(p: Pipe):
    """Concatenated FSILS face data: (nNo per face, glob, val(3,n) with val = 0)."""
    nn = np.array([len(f[1]) for f in p.faces], dtype=np.int32)
    glob = np.concatenate([f[1] for f in p.faces]).astype(np.int32)
    val = np.zeros((glob.shape[0], 3))
    return nn, glob, val
