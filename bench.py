#!/usr/bin/env python3
"""Benchmark of the fluid Newton step on one MI355X.

Workload (default): the seeded 1.2M-tet VMS pipe (n = 34: 1 179 120 tets,
209 475 nodes, 4 dof/node, 4x4 FP64 blocks), rho = 1.06, mu = 0.04, parabolic
inflow with peak 30, no-slip walls, zero-traction outlet, dt = 1e-3, one time
step of 3 Newton iterations; FSILS GMRES (Krylov dimension 50) with the
diagonal preconditioner to a 1e-6 relative FP64 residual (absolute tolerance
1e-30 so the relative test binds).

One bench "step" = one such time step from the same initial state (reset by
device-to-device copies, inside the timed region); `value` is seconds per
Newton step = elapsed / (3 * steps). Inputs are resident in HBM before timing.

Prints one JSON line with the metric, the roofline of the dominant kernel
(HIP events on the launching stream over the timed region) and the CPU
baseline (the single-core oracle restatement of the reference on a bounded
sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
KERNELS = ["spmv_bcsr", "spmv32", "d2f", "mdot", "mdot_fin", "gs_update", "div", "backsolve", "xupdate", "pc_w", "pc_val",
           "fluid_tet4", "pici", "picc", "fill", "wc_scale", "set_int", "spmv_tune"]


def kernel_table(lib):
    import ctypes as C
    out = {}
    for k in KERNELS:
        cnt, ms = C.c_int(), C.c_double()
        lib.svg_prof_get(k.encode(), C.byref(cnt), C.byref(ms))
        if cnt.value:
            out[k] = {"count": cnt.value, "total_ms": round(ms.value, 4), "avg_us": round(1e3 * ms.value / cnt.value, 3)}
    return out


def cpu_baseline(pipe, phys, gpu_hist, budget_s=30.0):
    """Single-core oracle on a bounded sample of the same workload: the full
    C2 assembly once, one preconditioning pass, and one 50-step GMRES cycle;
    the per-Newton-step time is assembled as
    T_asm + T_pc + (GMRES iterations of the GPU run per Newton step) * T_iter."""
    import numpy as np
    import oracle
    from svfsiplus_amd.workload import dirichlet_state
    Ao, Yo, An, Yn = dirichlet_state(pipe, phys)
    Ag, Yg = oracle.pici(phys, Ao, An, Yo, Yn)
    t0 = time.perf_counter()
    R, V = oracle.construct_fluid(pipe, phys, Ag, Yg)
    t_asm = time.perf_counter() - t0
    t0 = time.perf_counter()
    V2, R2, _ = oracle.precond_diag(pipe, V, R)
    t_pc = time.perf_counter() - t0
    sD = 50
    ls = oracle.ls_create(oracle.LS_GMRES, relTol=1e-30, absTol=1e-300, maxItr=1, dimKry=sD)
    t0 = time.perf_counter()
    oracle.fsils_solve(pipe, ls, R, V)
    t_solve1 = time.perf_counter() - t0 - t_pc
    t_iter = max(t_solve1, 1e-9) / max(ls.RI.itr, 1)
    itr_per_newton = float(np.mean([h["ls_itr"] for h in gpu_hist]))
    t_newton = t_asm + t_pc + itr_per_newton * t_iter
    return {
        "value": round(t_newton, 4),
        "unit": "s/Newton-step",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle (C restatement of the reference, 1 thread, -O2) on the same 1.2M-tet pipe: full "
                   f"assembly {t_asm:.2f}s, precond_diag {t_pc:.2f}s, one {ls.RI.itr}-SpMV GMRES cycle "
                   f"-> {1e3 * t_iter:.1f} ms/iteration x {itr_per_newton:.1f} iterations per Newton step "
                   f"(the GPU run's count)"),
        "host_cpu": platform.processor() or platform.machine(),
        "nproc": os.cpu_count(),
    }


def kernel_only(a, fn, pipe, lib):
    """SpMV launches on the first Newton system of the workload (assembled on the device),
    and a torch device copy of the SpMV's algorithmic byte count for counter calibration."""
    import ctypes as C
    import torch
    from svfsiplus_amd import _lib
    from svfsiplus_amd.newton import _stream, construct_fluid
    ph = fn.phys
    _lib.check(lib.svg_pici(pipe.nNo, ph.am, ph.af, fn.Ao.data_ptr(), fn.An.data_ptr(), fn.Yo.data_ptr(),
                            fn.Yn.data_ptr(), fn.Ag.data_ptr(), fn.Yg.data_ptr(), _stream()))
    fn.R.zero_()
    fn.Val.zero_()
    construct_fluid(fn.msh, ph, fn.Ag, fn.Yg, fn.R, fn.Val)
    U = torch.randn((pipe.nNo, 4), dtype=torch.float64, device="cuda")
    KU = torch.empty_like(U)
    nbytes = pipe.nnz * (16 * 8 + 4) + pipe.nNo * (4 + 4 * 8 + 4 * 8)
    src = torch.empty(nbytes // 2 // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    nv = C.c_int()
    _lib.check(lib.svg_spmv_num_variants(C.byref(nv)))
    variants = {}
    for v in range(nv.value):
        us = C.c_double()
        _lib.check(lib.svg_spmv_time_variant(fn.lhs.h, v, fn.Val.data_ptr(), U.data_ptr(), KU.data_ptr(), a.launches,
                                             C.byref(us), _stream()))
        variants[v] = round(us.value, 2)
    best = C.c_int()
    _lib.check(lib.svg_spmv_tune(fn.lhs.h, fn.Val.data_ptr(), C.byref(best), _stream()))
    lib.svg_prof_reset()
    lib.svg_prof_enable(1)
    for _ in range(a.launches):
        _lib.check(lib.svg_spar_mul_vv(fn.lhs.h, 4, fn.Val.data_ptr(), U.data_ptr(), KU.data_ptr(), _stream()))
    torch.cuda.synchronize()
    lib.svg_prof_enable(0)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(a.launches):
        dst.copy_(src)
    ev1.record()
    torch.cuda.synchronize()
    cnt, ms = C.c_int(), C.c_double()
    lib.svg_prof_get(b"spmv_bcsr", C.byref(cnt), C.byref(ms))
    avg = ms.value / cnt.value * 1e-3
    copy_avg = ev0.elapsed_time(ev1) / a.launches * 1e-3
    print(json.dumps({"kernel": "spmv_bcsr", "launches": cnt.value, "bytes_per_launch": nbytes,
                      "avg_us": round(avg * 1e6, 2), "achieved_GBs": round(nbytes / avg / 1e9, 1),
                      "variant_avg_us": variants, "tuned_variant": best.value,
                      "copy_bytes_per_launch": 2 * src.numel() * 8, "copy_avg_us": round(copy_avg * 1e6, 2),
                      "copy_GBs": round(2 * src.numel() * 8 / copy_avg / 1e9, 1)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=34, help="pipe cross-section cells (34 -> 1.2M tets)")
    ap.add_argument("--krylov", type=int, default=50)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-prof", action="store_true", help="time without per-kernel events")
    ap.add_argument("--kernel-only", choices=["spmv"], default=None,
                    help="launch only the block SpMV on the assembled system (for counter runs), "
                         "plus a device copy of the same byte count as the calibration kernel")
    ap.add_argument("--launches", type=int, default=20)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="fp64: the plain FSILS GMRES; mixed: FP32 block SpMV inside FP64-refined GMRES")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (the multi-GPU sharding is a design, not built); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from svfsiplus_amd import _lib, workload as W
    from svfsiplus_amd.newton import FluidNewton, ls_create

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    lib = _lib.load()
    t_setup = time.perf_counter()
    pipe = W.make_pipe(a.n)
    phys = W.Physics()
    ls = ls_create(_lib.LS_GMRES, relTol=a.tol, absTol=1e-30, maxItr=1000, dimKry=a.krylov)
    precision = _lib.PRECISION_MIXED if a.precision == "mixed" else _lib.PRECISION_FP64
    fn = FluidNewton(pipe, phys, ls, precision=precision)
    t_setup = time.perf_counter() - t_setup
    print(f"# setup {t_setup:.1f}s: nNo={pipe.nNo} nEl={pipe.nEl} nnzb={pipe.nnz} colours={fn.msh.nColors}",
          file=sys.stderr, flush=True)

    if a.kernel_only:
        kernel_only(a, fn, pipe, lib)
        return

    for w in range(a.warmup):
        fn.reset()
        h = fn.time_step()
        print(f"# warmup {w}: " + " ".join(f"[itr {x['itr']} ls {x['ls_itr']} iNorm {x['iNorm']:.3e}]" for x in h),
              file=sys.stderr, flush=True)

    lib.svg_prof_reset()
    lib.svg_prof_enable(0 if a.no_prof else 1)
    hists = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        fn.reset()
        hists.append([dict(x) for x in fn.time_step()])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    lib.svg_prof_enable(0)
    ktab = kernel_table(lib)
    for k, h in enumerate(hists):
        print(f"# step {k}: " + " ".join(f"[itr {x['itr']} ls {x['ls_itr']} iNorm {x['iNorm']:.3e} "
                                          f"fNorm {x['fNorm']:.3e}]" for x in h), file=sys.stderr, flush=True)
    n_newton = sum(len(h) for h in hists)
    elapsed = t1 - t0
    s_per_newton = elapsed / n_newton

    # Roofline of the dominant kernel: the 4x4 block SpMV.
    # Inside GMRES the SpMV is fused with the Gram-Schmidt dots of the Arnoldi
    # step, so its algorithmic bytes grow with the step index; the library sums
    # them per launch (132 B per block + 68 B per row + 32 B per row per basis
    # vector read).
    kname = "spmv32" if precision == _lib.PRECISION_MIXED else "spmv_bcsr"
    sp = ktab.get(kname, {"count": 0, "total_ms": 0.0})
    roof = None
    if sp["count"]:
        import ctypes as C
        tb = C.c_double()
        lib.svg_prof_get_bytes(kname.encode(), C.byref(tb))
        avg_s = sp["total_ms"] / sp["count"] * 1e-3
        bytes_per_launch = tb.value / sp["count"]
        ach = bytes_per_launch / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kname,
                "bytes_per_launch": round(bytes_per_launch), "avg_us": round(avg_s * 1e6, 2)}
    total_kernel_ms = sum(v["total_ms"] for v in ktab.values())

    cpu = None
    if not a.no_cpu:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(pipe, phys, [x for h in hists for x in h])

    out = {
        "metric": "Newton-step s at FP64 rel-res 1e-6, 1.2M-tet VMS pipe: learned-PC vs plain",
        "value": round(s_per_newton, 6),
        "unit": "s/Newton-step",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * elapsed / a.steps, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded pipe generator)",
        "config": {"workload": f"VMS pipe n={a.n}: {pipe.nEl} tets, {pipe.nNo} nodes, nnzb {pipe.nnz}; "
                               f"1 time step x 3 Newton iterations; GMRES({a.krylov}) diag PC rel-tol {a.tol:g}",
                   "solver": ("FSILS GMRES (gmres_v), diagonal preconditioner, FP64" if precision == _lib.PRECISION_FP64
                              else "GMRES with FP32 block SpMV, FP64 Gram-Schmidt/Givens, FP64 true-residual "
                                   "restarts, diagonal preconditioner"),
                   "method": "plain" if precision == _lib.PRECISION_FP64 else "mixed",
                   "newton_records": hists[-1]},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": ktab,
        "kernel_ms_per_newton": round(total_kernel_ms / n_newton, 3),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
