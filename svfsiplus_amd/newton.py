"""Host-side driver of the fluid Newton step on the device.

Mirrors the reference's Newton loop for one fluid equation
(Code/Source/solver/main.cpp:371-617):

    pici   (pic.cpp:486-574)          Ag, Yg from Ao/An, Yo/Yn
    ls_alloc                          R = 0, Val = 0
    construct_fluid (fluid.cpp:464)   element assembly          -> svg_construct_fluid
    set_bc_neu -> b_assem_neu_bc      outlet backflow term      -> svg_b_assem_neu_bc
    ls_solve -> fsils_solve (solve.cpp:50)                       -> svg_fsils_solve
    picc   (pic.cpp:74-330)           An -= R, Yn -= gam dt R, Newton convergence test

Device arrays are torch tensors (HBM storage and the stream); all arithmetic
runs in libsvgpu.so. There is no CPU fallback: without the library or a GPU
every call raises.
"""
from __future__ import annotations

import ctypes as C
import sys
from dataclasses import dataclass

import numpy as np

from . import _lib
from .workload import Physics, Pipe, dirichlet_faces, dirichlet_state, outlet_faces


def _ptr(t) -> int:
    return t.data_ptr()


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class DeviceLhs:
    """Device-resident FSILS lhs (fsils_lhs_create + fsils_bc_create). A face is
    (name, nodes) — a Dirichlet face, val = 0 — or (name, nodes, "neu", val(nNo, 3))
    — a Neumann face, coupled when fsils_solve is given a nonzero res for it."""

    def __init__(self, pipe: Pipe, faces=None, d_ien=None):
        """With d_ien (a device tensor holding pipe.IEN) the block pattern is
        built on the device (svg_lhs_create_dev) instead of taken from
        pipe.rowPtr / pipe.colPtr."""
        lib = _lib.load()
        faces = pipe.faces if faces is None else faces
        h = C.c_void_p()
        if d_ien is None:
            _lib.check(lib.svg_lhs_create(pipe.nNo, pipe.nnz, pipe.rowPtr.ctypes.data, pipe.colPtr.ctypes.data,
                                          len(faces), C.byref(h)))
            nnz = pipe.nnz
        else:
            _lib.check(lib.svg_lhs_create_dev(pipe.nNo, pipe.nEl, _ptr(d_ien), len(faces), C.byref(h), _stream()))
            c = C.c_int()
            _lib.check(lib.svg_lhs_nnz(h, C.byref(c)))
            nnz = c.value
        self.h = h
        self.nNo, self.nnz, self.nFaces = pipe.nNo, nnz, len(faces)
        for i, f in enumerate(faces):
            g = np.ascontiguousarray(f[1], dtype=np.int32)
            neu = len(f) > 2 and f[2] == "neu"
            val = (np.ascontiguousarray(f[3], dtype=np.float64).reshape(len(g), 3) if neu else np.zeros((len(g), 3)))
            _lib.check(lib.svg_lhs_bc_create(h, i, len(g), 3, _lib.BC_NEU if neu else _lib.BC_DIR, g.ctypes.data,
                                             val.ctypes.data))

    def free(self):
        if self.h:
            _lib.load().svg_lhs_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_arrays(pipe: Pipe):
    """pipe.IEN (int32) and pipe.x (float64) uploaded for the device setup."""
    import torch
    return (torch.as_tensor(np.ascontiguousarray(pipe.IEN, dtype=np.int32), device="cuda"),
            torch.as_tensor(np.ascontiguousarray(pipe.x, dtype=np.float64), device="cuda"))


class DeviceMesh:
    """Device-resident element data (coordinates, connectivity, colours, slot map)."""

    def __init__(self, lhs: DeviceLhs, pipe: Pipe, d_ien=None, d_x=None):
        """With d_ien and d_x (device tensors holding pipe.IEN and pipe.x) the
        colouring, colour order and slot map are built on the device
        (svg_msh_create_dev)."""
        lib = _lib.load()
        h = C.c_void_p()
        if d_ien is None:
            IEN = np.ascontiguousarray(pipe.IEN, dtype=np.int32)
            x = np.ascontiguousarray(pipe.x)
            _lib.check(lib.svg_msh_create(lhs.h, pipe.nEl, IEN.ctypes.data, pipe.nNo, x.ctypes.data, C.byref(h)))
        else:
            _lib.check(lib.svg_msh_create_dev(lhs.h, pipe.nEl, _ptr(d_ien), _ptr(d_x), C.byref(h), _stream()))
        self.h = h
        nc = C.c_int()
        _lib.check(lib.svg_msh_num_colors(h, C.byref(nc)))
        self.nColors = nc.value

    def free(self):
        if self.h:
            _lib.load().svg_msh_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceFace:
    """A TRI3 Neumann face of the fluid mesh on the device (faceType IEN, gE)."""

    def __init__(self, msh: DeviceMesh, faceIEN, gE):
        lib = _lib.load()
        h = C.c_void_p()
        fi = np.ascontiguousarray(faceIEN, dtype=np.int32)
        ge = np.ascontiguousarray(gE, dtype=np.int32)
        _lib.check(lib.svg_face_create(msh.h, fi.shape[0], fi.ctypes.data, ge.ctypes.data, C.byref(h)))
        self.h = h
        self.nFe = fi.shape[0]

    def free(self):
        if self.h:
            _lib.load().svg_face_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def b_assem_neu_bc(face: DeviceFace, phys: Physics, Yg, R, Val, hg=None):
    """Neumann-face assembly with backflow stabilisation (eq_assem.cpp:58-172)."""
    prm = phys.fluid_params()
    _lib.check(_lib.load().svg_b_assem_neu_bc(face.h, C.byref(prm), phys.beta, phys.qmTRI3,
                                              None if hg is None else _ptr(hg), _ptr(Yg), _ptr(R), _ptr(Val),
                                              _stream()))


def ls_create(ls_type: int, relTol=None, absTol=None, maxItr=None, dimKry=None, ns_gm=None, ns_cg=None) -> _lib.Ls:
    """fsils_ls_create defaults (ls.cpp:40-120) followed by the <LS> overrides
    of read_ls (read_files.cpp:2056-2090). ns_gm/ns_cg = (max_iterations, tolerance)."""
    ls = _lib.Ls()
    _lib.check(_lib.load().svg_ls_create(C.byref(ls), ls_type))
    if maxItr is not None:
        ls.RI.mItr = maxItr
    if relTol is not None:
        ls.RI.relTol = relTol
    if absTol is not None:
        ls.RI.absTol = absTol
    if dimKry is not None:
        ls.RI.sD = dimKry
    if ls_type == _lib.LS_NS:
        gm = ns_gm or (1000, 1e-2)
        cg = ns_cg or (1000, 1e-2)
        ls.GM.mItr, ls.GM.relTol = gm
        ls.CG.mItr, ls.CG.relTol = cg
        ls.GM.absTol = ls.RI.absTol
        ls.CG.absTol = ls.RI.absTol
        ls.GM.sD = ls.RI.sD
    return ls


def construct_fluid(msh: DeviceMesh, phys: Physics, Ag, Yg, R, Val, Bf=None):
    prm = phys.fluid_params()
    _lib.check(_lib.load().svg_construct_fluid(msh.h, C.byref(prm), _ptr(Ag), _ptr(Yg),
                                               None if Bf is None else _ptr(Bf), _ptr(R), _ptr(Val), _stream()))


def solve_opts(precision: int = _lib.PRECISION_FP64, lpc=None, fallback_window: int = 0,
               net_period: int = 1, coarse: int = 0, coarse_keep: bool = False) -> _lib.SolveOpts:
    """The svg_solve_opts of a solve: precision; the learned method when a
    network is given; coarse > 0 adds the two-level preconditioner with node
    aggregates of about `coarse` nodes (smoothed by block Jacobi, or by the
    network on its steps)."""
    opts = _lib.SolveOpts()
    opts.precision = precision
    if lpc is not None:
        opts.method = _lib.METHOD_LEARNED
        opts.lpc = lpc.h.value if hasattr(lpc.h, "value") else lpc.h
        opts.net_period = net_period
    if lpc is not None or coarse > 0:
        opts.fallback_window = fallback_window
        opts.coarse = coarse
        opts.coarse_keep = 1 if (coarse > 0 and coarse_keep) else 0
    return opts


def fsils_solve(lhs: DeviceLhs, ls: _lib.Ls, dof: int, Ri, Val, prec: int = _lib.PREC_FSILS, incL=None, res=None,
                precision: int = _lib.PRECISION_FP64, lpc=None, fallback_window: int = 0, net_period: int = 1,
                info: dict | None = None, coarse: int = 0, coarse_keep: bool = False):
    """Reference signature: fsils_solve(lhs, ls, dof, Ri, Val, prec, incL, res) (solve.cpp:50).
    precision = PRECISION_MIXED selects the FP32-SpMV / FP64-refined GMRES; a
    network `lpc` (LearnedPC) or coarse > 0 (the two-level preconditioner,
    aggregates of about `coarse` nodes) selects the flexible GMRES
    (info["fallback_itr"] = iteration count at the fallback, -1 if none);
    coarse_keep reuses the coarse inverse of the lhs's last two-level set-up."""
    nf = lhs.nFaces
    inc = np.ones(max(1, nf), dtype=np.int32) if incL is None else np.ascontiguousarray(incL, dtype=np.int32)
    rs = None if res is None else np.ascontiguousarray(res, dtype=np.float64)
    opts = solve_opts(precision, lpc, fallback_window, net_period, coarse, coarse_keep)
    _lib.check(_lib.load().svg_fsils_solve_ex(lhs.h, C.byref(ls), dof, _ptr(Ri), _ptr(Val), prec, nf,
                                              inc.ctypes.data, None if rs is None else rs.ctypes.data, C.byref(opts),
                                              _stream()))
    if info is not None:
        info["fallback_itr"] = opts.fallback_itr
    return ls


def is_zero(v: float) -> bool:
    """utils::is_zero (utils.cpp:170-189) with value2 = 0."""
    a = abs(v)
    eps = sys.float_info.epsilon
    return a / max(a, eps) < 10.0 * eps


@dataclass
class EqState:
    """The convergence bookkeeping of eqType used by picc."""
    minItr: int = 3
    maxItr: int = 3
    tol: float = 0.5
    itr: int = 0
    iNorm: float = 0.0
    pNorm: float = float(np.finfo(np.float64).max)
    ok: bool = False


class FluidNewton:
    """One fluid equation on one pipe, all state resident in HBM."""

    def __init__(self, pipe: Pipe, phys: Physics, ls: _lib.Ls, eq: EqState | None = None, device="cuda",
                 precision: int = _lib.PRECISION_FP64, lpc=None, outlet: bool = True, net_period: int = 1,
                 device_setup: bool = False, coarse: int = 0, coarse_keep: bool = False):
        """device_setup: build the block pattern, colouring and slot map on the
        device (svg_lhs_create_dev, svg_msh_create_dev) from IEN and x;
        coarse_keep: the Newton iterations after the first of a time step keep
        the coarse inverse of the first (the two-level preconditioner)."""
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("FluidNewton needs a GPU (the HIP library has no CPU path)")
        self.torch = torch
        self.pipe, self.phys, self.ls = pipe, phys, ls
        self.precision = precision
        self.lpc = lpc
        self.net_period = net_period
        self.coarse = coarse
        self.coarse_keep = coarse_keep
        self.eq = eq or EqState()
        if device_setup:
            d_ien, d_x = device_arrays(pipe)
            self.lhs = DeviceLhs(pipe, d_ien=d_ien)
            self.msh = DeviceMesh(self.lhs, pipe, d_ien=d_ien, d_x=d_x)
        else:
            self.lhs = DeviceLhs(pipe)
            self.msh = DeviceMesh(self.lhs, pipe)
        # the zero-traction outlet is a Neumann face: backflow stabilisation (b_assem_neu_bc)
        self.outlet = None
        if outlet:
            fi, ge, _ = outlet_faces(pipe)
            if fi.shape[0]:
                self.outlet = DeviceFace(self.msh, fi, ge)
        dev = torch.device(device)
        f64 = torch.float64
        n = pipe.nNo
        Ao, Yo, An, Yn = dirichlet_state(pipe, phys)
        self.Ao = torch.tensor(Ao, dtype=f64, device=dev)
        self.Yo = torch.tensor(Yo, dtype=f64, device=dev)
        self.An = torch.tensor(An, dtype=f64, device=dev)
        self.Yn = torch.tensor(Yn, dtype=f64, device=dev)
        self.Ag = torch.empty((n, 4), dtype=f64, device=dev)
        self.Yg = torch.empty((n, 4), dtype=f64, device=dev)
        self.R = torch.empty((n, 4), dtype=f64, device=dev)
        self._init = [t.clone() for t in (self.Ao, self.Yo, self.An, self.Yn)]
        self.Val = torch.empty((self.lhs.nnz, 16), dtype=f64, device=dev)
        self.history = []
        # strong Dirichlet faces for set_bc_dir at later time steps (node ids, gx nV)
        self._dir = [(torch.tensor(g, dtype=torch.int32, device=dev), torch.tensor(v, dtype=f64, device=dev))
                     for g, v in dirichlet_faces(pipe, phys)]
        self.step = 0

    def set_physics(self, phys: Physics):
        """Another (rho, mu, inflow peak, ...) on the same mesh: the initial
        state and the strong Dirichlet values follow the new inflow; the
        pattern, colouring and device buffers are kept."""
        torch = self.torch
        self.phys = phys
        init = dirichlet_state(self.pipe, phys)
        self._init = [torch.tensor(a, dtype=torch.float64, device=self.Ao.device) for a in init]
        self._dir = [(torch.tensor(g, dtype=torch.int32, device=self.Ao.device),
                      torch.tensor(v, dtype=torch.float64, device=self.Ao.device))
                     for g, v in dirichlet_faces(self.pipe, phys)]
        self.reset()

    def reset(self):
        """Back to the initial state of the time step (device-to-device copies)."""
        for t, a in zip((self.Ao, self.Yo, self.An, self.Yn), self._init):
            t.copy_(a)
        self.eq = EqState(self.eq.minItr, self.eq.maxItr, self.eq.tol)
        self.history = []
        self.step = 0

    def newton_iteration(self):
        lib, ph, eq, ls = _lib.load(), self.phys, self.eq, self.ls
        s = _stream()
        eq.itr += 1
        _lib.check(lib.svg_pici(self.pipe.nNo, ph.am, ph.af, _ptr(self.Ao), _ptr(self.An), _ptr(self.Yo),
                                _ptr(self.Yn), _ptr(self.Ag), _ptr(self.Yg), s))
        _lib.check(lib.svg_fill(_ptr(self.R), self.R.numel(), 0.0, s))
        _lib.check(lib.svg_fill(_ptr(self.Val), self.Val.numel(), 0.0, s))
        construct_fluid(self.msh, ph, self.Ag, self.Yg, self.R, self.Val)
        if self.outlet is not None:
            b_assem_neu_bc(self.outlet, ph, self.Yg, self.R, self.Val)
        info = {}
        fsils_solve(self.lhs, ls, 4, self.R, self.Val, precision=self.precision, lpc=self.lpc,
                    net_period=self.net_period, info=info, coarse=self.coarse,
                    coarse_keep=self.coarse_keep and eq.itr > 1)
        _lib.check(lib.svg_picc_update(self.pipe.nNo, ph.gam, ph.dt, _ptr(self.R), _ptr(self.An), _ptr(self.Yn), s))
        # Newton convergence, pic.cpp:236-281
        eps = np.finfo(np.float64).eps
        if is_zero(ls.RI.iNorm):
            ls.RI.iNorm = eps
        if is_zero(eq.iNorm):
            eq.iNorm = ls.RI.iNorm
        if eq.itr == 1:
            eq.pNorm = ls.RI.iNorm / eq.iNorm
        r1 = ls.RI.iNorm / eq.iNorm
        l1 = eq.itr >= eq.maxItr
        l2 = r1 <= eq.tol
        l3 = r1 <= eq.tol * eq.pNorm
        l4 = eq.itr >= eq.minItr
        if l1 or ((l2 or l3) and l4):
            eq.ok = True
        self.history.append(dict(itr=eq.itr, ls_itr=ls.RI.itr, iNorm=ls.RI.iNorm, fNorm=ls.RI.fNorm, suc=ls.RI.suc,
                                 r1=r1, fallback_itr=info.get("fallback_itr", -1)))
        return eq.ok

    def time_step(self):
        """One time step: Newton iterations until picc reports convergence."""
        self.eq.itr = 0
        self.eq.ok = False
        while not self.newton_iteration():
            pass
        return self.history

    def advance(self):
        """To the next time step on the device: Ao = An, Yo = Yn (main.cpp:754-755),
        the predictor (picp, pic.cpp:591-715) and the strong Dirichlet values
        of every face (set_bc_dir, set_bc.cpp:794-1063; steady: dirA = 0, dirY = 1)."""
        self.Ao.copy_(self.An)
        self.Yo.copy_(self.Yn)
        self._predict()
        self.step += 1

    def _predict(self):
        lib, s = _lib.load(), _stream()
        _lib.check(lib.svg_picp(self.pipe.nNo, self.phys.gam, _ptr(self.Ao), _ptr(self.Yo), _ptr(self.An),
                                _ptr(self.Yn), s))
        for gN, gxn in self._dir:
            _lib.check(lib.svg_set_bc_dir(gN.shape[0], _ptr(gN), 4, 0, 3, 0, _ptr(gxn), 0.0, 1.0, _ptr(self.An),
                                          _ptr(self.Yn), s))
        self.history = []

    def restart_stamp(self) -> tuple:
        """The stamp of this run's restart records (one rank, one equation, one mesh, no coupled-BC state)."""
        from . import restart
        return restart.make_stamp(self.pipe.nNo, tDof=4)

    def save_restart(self, stFileName: str, replace: bool = False, cpu_time: float = 0.0) -> str:
        """write_restart (output.cpp:202-360) after time step `step + 1`: Yn, An
        and eq.iNorm, the file named for the time step with a `_last.bin` link."""
        from . import restart
        self.torch.cuda.synchronize()
        cTS = self.step + 1
        r = restart.Restart(self.restart_stamp(), cTS, cTS * self.phys.dt, cpu_time,
                            np.array([self.eq.iNorm]), self.Yn.cpu().numpy(), self.An.cpu().numpy())
        return restart.save(stFileName, r, replace)

    def load_restart(self, fname: str):
        """init_from_bin (initialize.cpp:78-241): Yo, Ao and eq.iNorm from the
        record, then the predictor and Dirichlet values of the next time step
        (cTS + 1) on the device, as `advance` leaves them."""
        from . import restart
        r = restart.read(fname, self.restart_stamp())
        self.Yo.copy_(self.torch.from_numpy(r.Yn))
        self.Ao.copy_(self.torch.from_numpy(r.An))
        self.eq = EqState(self.eq.minItr, self.eq.maxItr, self.eq.tol, iNorm=float(r.iNorm[0]))
        self.step = r.cTS
        self._predict()
        return r

    def time_steps(self, nsteps: int):
        """nsteps time steps from the current state; returns the Newton history of each."""
        out = []
        for k in range(nsteps):
            if k:
                self.advance()
            out.append(list(self.time_step()))
        return out
