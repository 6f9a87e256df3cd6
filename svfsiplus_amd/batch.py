"""Batched Newton steps over many independent meshes on one GPU.

The meshes are concatenated into one block-diagonal union (nodes, elements,
block-CSR pattern and Dirichlet faces offset per mesh), so the assembly
(construct_fluid + the outlets' b_assem_neu_bc), pici and picc run as single
launches over all meshes, and the linear solve is the batched GMRES of the
library (per-mesh Krylov spaces and convergence). A Newton iteration is
svg_batch_newton_iteration: three hipGraphs captured once and replayed (pici,
assembly, preconditioner and the start of GMRES; one restart cycle of every
mesh, replayed until all have converged; the solutions and the corrector).
Each mesh's solution is the one a solve of that mesh alone gives.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib
from .newton import FluidNewton, _ptr, _stream, b_assem_neu_bc, construct_fluid, ls_create, solve_opts
from .workload import Physics, Pipe, make_bifurcation


def union(meshes: list[Pipe]) -> tuple[Pipe, np.ndarray]:
    """Block-diagonal union of the meshes and the node offsets (len + 1)."""
    noff = np.zeros(len(meshes) + 1, dtype=np.int64)
    noff[1:] = np.cumsum([m.nNo for m in meshes])
    boff = np.zeros(len(meshes) + 1, dtype=np.int64)
    boff[1:] = np.cumsum([m.nnz for m in meshes])
    if noff[-1] > 2**31 - 1 or boff[-1] > 2**31 - 1:
        raise ValueError("batch too large for 32-bit indices")
    if any(m.radius != meshes[0].radius for m in meshes):
        raise ValueError("the meshes of a batch share one inlet radius (the inflow profile is set per union)")
    x = np.concatenate([m.x for m in meshes])
    IEN = np.concatenate([m.IEN + o for m, o in zip(meshes, noff)]).astype(np.int32)
    flags = np.concatenate([m.flags for m in meshes]).astype(np.int32)
    rowPtr = np.concatenate([m.rowPtr[:-1] + b for m, b in zip(meshes, boff)] + [np.array([boff[-1]])]).astype(np.int32)
    colPtr = np.concatenate([m.colPtr + o for m, o in zip(meshes, noff)]).astype(np.int32)
    u = Pipe(meshes[0].n, meshes[0].nz, np.ascontiguousarray(x), np.ascontiguousarray(IEN), flags, rowPtr, colPtr,
             meshes[0].radius, meshes[0].length)
    u.faces = [(name, np.concatenate([dict(m.faces)[name] + o for m, o in zip(meshes, noff)]).astype(np.int32))
               for name, _ in meshes[0].faces]
    return u, noff.astype(np.int32)


class BatchNewton:
    """Newton iterations of the fluid equation on every mesh of a batch at once."""

    def __init__(self, meshes: list[Pipe], phys: Physics, sD: int = 50, rel_tol: float = 1e-6,
                 precision: int = _lib.PRECISION_FP64, max_restarts: int = 1000, lpc=None, net_period: int = 5,
                 fallback_window: int = 20, device_setup: bool = False, coarse: int = 0):
        self.meshes = meshes
        self.nsys = len(meshes)
        self.pipe, self.noff = union(meshes)
        self.phys = phys
        self.precision = precision
        self.lpc = lpc
        # the learned preconditioner (flexible GMRES, per-mesh fallback) when a network is given, or
        # the two-level preconditioner with aggregates of about `coarse` nodes (one coarse block per mesh)
        self.coarse = coarse
        self.opts = solve_opts(precision, lpc, fallback_window, net_period, coarse)
        self.fn = FluidNewton(self.pipe, phys, ls_create(_lib.LS_GMRES, relTol=rel_tol, absTol=1e-30,
                                                         maxItr=max_restarts, dimKry=sD), device_setup=device_setup)
        h = C.c_void_p()
        _lib.check(_lib.load().svg_batch_create(self.fn.lhs.h, self.nsys, self.noff.ctypes.data, sD, C.byref(h)))
        self.h = h
        self.proto = _lib.SubLs()
        self.proto.mItr, self.proto.sD = max_restarts, sD
        self.proto.absTol, self.proto.relTol = 1e-30, rel_tol
        self.out = (_lib.SubLs * self.nsys)()

    def reset(self):
        self.fn.reset()

    def newton_iteration(self, graph: bool = True):
        """pici, assembly of every mesh, the batched fsils_solve, picc update.
        Returns the per-mesh linear solve records (itr, iNorm, fNorm, suc).
        graph=False enqueues the same launches one by one (the eager form)."""
        lib, fn, ph = _lib.load(), self.fn, self.phys
        s = _stream()
        if graph:
            prm = ph.fluid_params()
            _lib.check(lib.svg_batch_newton_iteration(
                self.h, fn.msh.h, None if fn.outlet is None else fn.outlet.h, C.byref(prm), ph.beta, ph.qmTRI3,
                _ptr(fn.Ao), _ptr(fn.An), _ptr(fn.Yo), _ptr(fn.Yn), _ptr(fn.Ag), _ptr(fn.Yg), _ptr(fn.R),
                _ptr(fn.Val), C.byref(self.proto), C.byref(self.opts), self.out, s))
            return self._records()
        n = self.pipe.nNo
        _lib.check(lib.svg_pici(n, ph.am, ph.af, _ptr(fn.Ao), _ptr(fn.An), _ptr(fn.Yo), _ptr(fn.Yn), _ptr(fn.Ag),
                                _ptr(fn.Yg), s))
        _lib.check(lib.svg_fill(_ptr(fn.R), fn.R.numel(), 0.0, s))
        _lib.check(lib.svg_fill(_ptr(fn.Val), fn.Val.numel(), 0.0, s))
        construct_fluid(fn.msh, ph, fn.Ag, fn.Yg, fn.R, fn.Val)
        if fn.outlet is not None:
            b_assem_neu_bc(fn.outlet, ph, fn.Yg, fn.R, fn.Val)
        _lib.check(lib.svg_batch_fsils_solve_ex(self.h, C.byref(self.proto), _ptr(fn.R), _ptr(fn.Val),
                                                C.byref(self.opts), self.out, s))
        _lib.check(lib.svg_picc_update(n, ph.gam, ph.dt, _ptr(fn.R), _ptr(fn.An), _ptr(fn.Yn), s))
        return self._records()

    def _records(self):
        recs = [dict(itr=o.itr, iNorm=o.iNorm, fNorm=o.fNorm, suc=o.suc) for o in self.out]
        if self.lpc is not None or self.coarse > 0:
            fb = np.empty(self.nsys, dtype=np.int32)
            _lib.check(_lib.load().svg_batch_fallback_itr(self.h, fb.ctypes.data))
            for r, f in zip(recs, fb):
                r["fallback_itr"] = int(f)
        return recs

    def replays(self) -> int:
        """Restart-cycle graph replays (host synchronisations) since the batch was created."""
        n = C.c_longlong()
        _lib.check(_lib.load().svg_batch_replays(self.h, C.byref(n)))
        return n.value

    def time_step(self, newton_iterations: int = 3, graph: bool = True):
        return [self.newton_iteration(graph) for _ in range(newton_iterations)]

    def free(self):
        if self.h:
            _lib.load().svg_batch_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def make_batch(specs) -> list[Pipe]:
    return [make_bifurcation(**sp) for sp in specs]
