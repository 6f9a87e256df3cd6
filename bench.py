#!/usr/bin/env python3
"""Benchmark of the fluid Newton step on one MI355X.

Workload (default): the seeded 1.2M-tet VMS pipe (n = 34: 1 179 120 tets,
209 475 nodes, 4 dof/node, 4x4 FP64 blocks), rho = 1.06, mu = 0.04, parabolic
inflow with peak 30, no-slip walls, zero-traction outlet, dt = 1e-3, one time
step of 3 Newton iterations; GMRES (Krylov dimension 50) to a 1e-6 relative
FP64 residual (absolute tolerance 1e-30 so the relative test binds).

Methods: "plain" = the reference's FSILS GMRES with the diagonal
preconditioner in FP64; "twolevel" = flexible GMRES with the FP32 block SpMV
(FP64 Gram-Schmidt, Givens and true-residual restarts) preconditioned by the
two-level preconditioner (pressure aggregation coarse level + block Jacobi);
"learned" = the same with the learned network as the smoother on every
net-period-th Arnoldi step (fallback: to the two-level preconditioner with
block Jacobi, then to the diagonal one); "mixed" = plain GMRES with the FP32
SpMV. The default, "compare", times plain, twolevel and learned back to back:
`value` is the learned method's seconds per Newton step, `plain`, `twolevel`,
`speedup_vs_plain` and `network_gain` stand beside it. The network is loaded
from svfsiplus_amd/data/lpc_smoother.npz (trained on the GPU by
`python -m svfsiplus_amd.lpc --coarse 128`), or trained here within a bounded
budget (outside the timed region) when that file is absent. On the 1.2M-tet
pipe twolevel and learned use aggregates of about 64 nodes and learned applies
the network on every 8th step (the fastest of a sweep over 32-256 nodes and
periods 5-16); the stenosed 4.2M-tet leg uses 128 nodes (64 would give more
aggregates than the dense coarse block allows) and every 5th step.

Further legs of the default run: the mixed-precision draws of viscosity and
inflow (diagonal preconditioner, C3), the 4.2M-tet stenosed pipe (C4: twolevel
and learned timed, plain FP64 on its first Newton system bounded by a restart
budget), the CPU baseline and the batched bifurcations (C5).

One bench "step" = one time step from the same initial state (reset by
device-to-device copies, inside the timed region); seconds per Newton step =
elapsed / (Newton iterations). Inputs are resident in HBM before timing.

Prints one JSON line with the metric, the roofline of the dominant kernel of
the headline run (HIP events on the launching stream over the timed region)
and the CPU baseline (the single-core oracle restatement of the reference on
a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json

import numpy as np
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
KERNELS = ["spmv_bcsr", "spmv32", "d2f", "mdot", "mdot_cb", "mdot_fin", "gs_update", "gs_update_cb", "zcopy", "div",
           "backsolve", "xupdate", "xupdate_cb",
           "pc_w", "pc_val", "fluid_elem", "fluid_gather", "b_fluid", "pici", "picc", "fill", "wc_scale", "set_int", "spmv_tune",
           "spmv32_tune", "lpc_prep", "lpc_encode", "lpc_msg", "lpc_dense", "lpc_layer", "add_bc_mul", "face_valM",
           "rsub", "cl_ap", "cl_dinv", "cl_acinit", "cl_ac", "gj_pivot", "gj_panels", "gj_update", "cl_d2f",
           "cl_restrict", "cl_gemv", "cl_apply", "cl_addc", "rearm"]
# kernels whose launches carry algorithmic byte counts (roofline candidates)
HBM_KERNELS = ["spmv_bcsr", "spmv32", "lpc_msg", "lpc_dense", "lpc_layer", "lpc_encode"]
ROOFLINE_KERNELS = ["spmv_bcsr", "spmv32"]  # timed with events inside the timed region
# the kernels of one Arnoldi step of the GMRES (SpMV, Gram-Schmidt dots, finaliser, update), FP64 / FP32 basis
ARNOLDI_KERNELS = {"fp64": ["spmv_bcsr", "mdot", "mdot_fin", "gs_update"],
                   "fp32": ["spmv32", "mdot_cb", "mdot_fin", "gs_update_cb"]}
PROF_EVERY = 16  # ... one launch in 16 of each (events between launches cost time)
DEFAULT_WEIGHTS = os.path.join(ROOT, "svfsiplus_amd", "data", "lpc_weights.npz")     # the network alone
DEFAULT_SMOOTHER = os.path.join(ROOT, "svfsiplus_amd", "data", "lpc_smoother.npz")   # the two-level smoother
# HBM bytes per launch on the 1.2M-tet pipe (n = 34) for each launch variant of the two SpMV kernels
# (the autotuners' variant indices), measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE on
# `bench.py --kernel-only spmv32` (which times every variant of both kernels), each scaled by the same
# counter's reading on a device copy of known size (FETCH_SIZE reports half of a copy's reads,
# WRITE_SIZE all of its writes).
MEASURED_TRAFFIC_N34 = {
    "spmv_bcsr": {0: 477.6e6, 1: 461.4e6, 2: 479.3e6, 3: 434.6e6, 4: 425.5e6, 5: 444.7e6, 6: 514.0e6, 7: 480.8e6,
                  8: 478.4e6, 9: 450.3e6, 10: 459.9e6, 11: 429.0e6},
    "spmv32": {0: 287.1e6, 1: 286.1e6, 2: 269.9e6, 3: 258.4e6, 4: 245.4e6, 5: 246.0e6},
}


def kernel_table(lib):
    import ctypes as C
    out = {}
    for k in KERNELS:
        cnt, ms = C.c_int(), C.c_double()
        lib.svg_prof_get(k.encode(), C.byref(cnt), C.byref(ms))
        if cnt.value:
            out[k] = {"count": cnt.value, "total_ms": round(ms.value, 4), "avg_us": round(1e3 * ms.value / cnt.value, 3)}
    return out


def host_cpu():
    """CPU model and clock of this host (from /proc/cpuinfo)."""
    model, mhz = platform.processor() or platform.machine(), None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k == "model name" and model in ("", "x86_64", platform.processor()):
                    model = v.strip()
                elif k == "cpu MHz" and mhz is None:
                    mhz = float(v)
    except OSError:
        pass
    return {"model": model, "mhz": mhz, "nproc": os.cpu_count()}


def cpu_baseline(pipe, phys, gpu_hist, cycles=4):
    """The reference's own CPU path, timed on this host: its construct_fluid
    (oracle/_ref, compiled from the reference's sources) over every element of
    the workload (measured), then its fsils_solve on that assembled system for
    `cycles` GMRES(50) restart cycles with the diagonal preconditioner
    (measured); the Newton step is the measured assembly plus the measured
    per-iteration solve time x the plain GPU run's GMRES iterations per Newton
    step (that product is the extrapolated part). Falls back to the oracle
    restatement ("port") where the reference build is not present."""
    import numpy as np
    import oracle
    from svfsiplus_amd.workload import dirichlet_state
    Ao, Yo, An, Yn = dirichlet_state(pipe, phys)
    Ag, Yg = oracle.pici(phys, Ao, An, Yo, Yn)
    itr_per_newton = float(np.mean([h["ls_itr"] for h in gpu_hist]))
    sD = 50
    cpu = host_cpu()
    kind, asm, solve = "port", oracle.construct_fluid, None
    if oracle.ref_available():
        kind, asm = "reference", oracle.ref_construct_fluid
    t0 = time.perf_counter()
    R, V = asm(pipe, phys, Ag, Yg)
    t_asm = time.perf_counter() - t0
    ls = oracle.ls_create(oracle.LS_GMRES, relTol=1e-30, absTol=1e-300, maxItr=cycles, dimKry=sD)
    t0 = time.perf_counter()
    if kind == "reference":
        _, _, rec = oracle.ref_fsils_solve(pipe, oracle.LS_GMRES, oracle.ls_params(ls), R, V)
        itr = int(rec["itr"])
    else:
        oracle.fsils_solve(pipe, ls, R, V)
        itr = int(ls.RI.itr)
    t_solve = time.perf_counter() - t0
    t_iter = t_solve / max(itr, 1)
    t_newton = t_asm + itr_per_newton * t_iter
    who = ("the reference's construct_fluid and fsils_solve (compiled from its sources)" if kind == "reference"
           else "the oracle (C restatement of the reference)")
    return {"value": round(t_newton, 3), "unit": "s/Newton-step", "cores": 1, "kind": kind,
            "sample": (f"{who}, 1 thread, on the same pipe ({pipe.nEl} tets, {pipe.nNo} nodes): assembly of every "
                       f"element measured ({t_asm:.2f}s); diagonal preconditioner + {itr} GMRES iterations "
                       f"({cycles} restart cycles of {sD}) measured ({t_solve:.2f}s, {1e3 * t_iter:.2f} ms/iteration); "
                       f"the solve is extrapolated to the plain GPU run's {itr_per_newton:.1f} iterations per "
                       f"Newton step"),
            "assembly_s": round(t_asm, 3), "solve_s_per_iteration": round(t_iter, 6),
            "solve_iterations_measured": itr, "solve_extrapolated": True,
            "host_cpu": cpu["model"], "host_mhz": cpu["mhz"], "nproc": cpu["nproc"]}


def kernel_only_lpc(a, fn, pipe, lib):
    """Applications of the learned preconditioner alone (encoder, 3 x message
    gather + dense layer, decoder) on the first Newton system of the workload,
    diagonally scaled as fsils_solve hands it to GMRES (for counter runs)."""
    import torch
    from svfsiplus_amd import _lib, lpc as L
    from svfsiplus_amd.newton import _stream, construct_fluid
    ph = fn.phys
    s = _stream()
    _lib.check(lib.svg_pici(pipe.nNo, ph.am, ph.af, fn.Ao.data_ptr(), fn.An.data_ptr(), fn.Yo.data_ptr(),
                            fn.Yn.data_ptr(), fn.Ag.data_ptr(), fn.Yg.data_ptr(), s))
    fn.R.zero_()
    fn.Val.zero_()
    construct_fluid(fn.msh, ph, fn.Ag, fn.Yg, fn.R, fn.Val)
    Wd = torch.empty((pipe.nNo, 4), dtype=torch.float64, device="cuda")
    _lib.check(lib.svg_precond_diag(fn.lhs.h, 4, fn.Val.data_ptr(), fn.R.data_ptr(), Wd.data_ptr(), s))
    _lib.check(lib.svg_lpc_prepare(fn.lhs.h, fn.Val.data_ptr(), s))
    net = L.LearnedPC.load(a.lpc_weights or (DEFAULT_SMOOTHER if a.coarse > 0 else DEFAULT_WEIGHTS))
    V = torch.randn((pipe.nNo, 4), dtype=torch.float64, device="cuda")
    V /= torch.linalg.vector_norm(V)
    Z = torch.empty_like(V)
    _lib.check(lib.svg_lpc_apply(net.h, fn.lhs.h, 1, V.data_ptr(), Z.data_ptr(), s))  # buffers and variant choice
    torch.cuda.synchronize()
    lib.svg_prof_reset()
    lib.svg_prof_enable(1)
    for _ in range(a.launches):
        _lib.check(lib.svg_lpc_apply(net.h, fn.lhs.h, 1, V.data_ptr(), Z.data_ptr(), s))
    torch.cuda.synchronize()
    lib.svg_prof_enable(0)
    tab = _bytes(lib, kernel_table(lib))
    print(json.dumps({"kernel": "lpc", "applications": a.launches, "nNo": pipe.nNo, "nnzb": pipe.nnz,
                      "kernels": {k: v for k, v in tab.items() if k.startswith("lpc_")}}))


def kernel_only_assembly(a, fn, pipe, lib):
    """construct_fluid alone (the element kernel and the element-order gather)
    on the first Newton iteration's state of the workload, a.launches times
    (for counter runs), with the per-kernel event table."""
    import torch
    from svfsiplus_amd import _lib
    from svfsiplus_amd.newton import _stream, construct_fluid
    ph = fn.phys
    _lib.check(lib.svg_pici(pipe.nNo, ph.am, ph.af, fn.Ao.data_ptr(), fn.An.data_ptr(), fn.Yo.data_ptr(),
                            fn.Yn.data_ptr(), fn.Ag.data_ptr(), fn.Yg.data_ptr(), _stream()))
    construct_fluid(fn.msh, ph, fn.Ag, fn.Yg, fn.R, fn.Val)
    torch.cuda.synchronize()
    lib.svg_prof_reset()
    lib.svg_prof_enable(1)
    for _ in range(a.launches):
        fn.R.zero_()
        fn.Val.zero_()
        construct_fluid(fn.msh, ph, fn.Ag, fn.Yg, fn.R, fn.Val)
    torch.cuda.synchronize()
    lib.svg_prof_enable(0)
    tab = kernel_table(lib)
    print(json.dumps({"kernel": "assembly", "launches": a.launches, "nEl": pipe.nEl, "nNo": pipe.nNo,
                      "nnzb": pipe.nnz,
                      "kernels": {k: v for k, v in tab.items() if k in ("fluid_elem", "fluid_gather")}}))


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
    kname = "spmv_bcsr"
    if a.kernel_only == "spmv32":
        # the mixed-precision solve's FP32 SpMV on FP32 copies of the same matrix and vector
        kname = "spmv32"
        K32 = fn.Val.to(torch.float32)
        scratch = torch.empty((pipe.nNo, 4), dtype=torch.float32, device="cuda")
        _lib.check(lib.svg_spmv32_tune(fn.lhs.h, K32.data_ptr(), scratch.data_ptr(), KU.data_ptr(), None, _stream()))
        U32 = U.to(torch.float32)
        nbytes = pipe.nnz * (16 * 4 + 4) + pipe.nNo * (4 + 4 * 4 + 4 * 8)
        src = torch.empty(nbytes // 2 // 8, dtype=torch.float64, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(a.launches):
            _lib.check(lib.svg_spmv32(fn.lhs.h, K32.data_ptr(), U32.data_ptr(), KU.data_ptr(), _stream()))
    else:
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
    lib.svg_prof_get(kname.encode(), C.byref(cnt), C.byref(ms))
    avg = ms.value / cnt.value * 1e-3
    copy_avg = ev0.elapsed_time(ev1) / a.launches * 1e-3
    print(json.dumps({"kernel": kname, "launches": cnt.value, "bytes_per_launch": nbytes,
                      "avg_us": round(avg * 1e6, 2), "achieved_GBs": round(nbytes / avg / 1e9, 1),
                      "variant_avg_us": variants, "tuned_variant": best.value,
                      "copy_bytes_per_launch": 2 * src.numel() * 8, "copy_avg_us": round(copy_avg * 1e6, 2),
                      "copy_GBs": round(2 * src.numel() * 8 / copy_avg / 1e9, 1)}))


def _bytes(lib, ktab):
    import ctypes as C
    for k in HBM_KERNELS + ARNOLDI_KERNELS["fp64"] + ARNOLDI_KERNELS["fp32"]:
        if k in ktab:
            tb = C.c_double()
            lib.svg_prof_get_bytes(k.encode(), C.byref(tb))
            ktab[k]["bytes"] = tb.value
    return ktab


def spmv_variants(lib, lhs):
    """The SpMV launch variants the library's autotuners picked for this lhs."""
    import ctypes as C
    v64, v32 = C.c_int(), C.c_int()
    lib.svg_spmv_variants(lhs.h, C.byref(v64), C.byref(v32))
    return {"fp64": v64.value, "fp32": v32.value}


def timed_run(fn, lib, steps, warmup, prof, label):
    """Warmup (every named kernel timed with events: the per-kernel table),
    then `steps` time steps timed with a device synchronisation on both sides,
    in which events bracket only one launch in PROF_EVERY of the block SpMV
    kernels the roofline reports on (events between launches cost time;
    prof=False: none at all)."""
    import torch
    warm_tab = {}
    for w in range(warmup):
        lib.svg_prof_reset()
        lib.svg_prof_only(None)
        lib.svg_prof_enable(1 if prof else 0)
        fn.reset()
        h = fn.time_step()
        lib.svg_prof_enable(0)
        warm_tab = _bytes(lib, kernel_table(lib))
        print(f"# {label} warmup {w}: " + " ".join(f"[itr {x['itr']} ls {x['ls_itr']} iNorm {x['iNorm']:.3e}]"
                                                 for x in h), file=sys.stderr, flush=True)
    lib.svg_prof_reset()
    lib.svg_prof_only(",".join(ROOFLINE_KERNELS).encode())
    lib.svg_prof_sample(PROF_EVERY)
    lib.svg_prof_enable(1 if prof else 0)
    hists = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn.reset()
        hists.append([dict(x) for x in fn.time_step()])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    lib.svg_prof_enable(0)
    lib.svg_prof_only(None)
    lib.svg_prof_sample(1)
    ktab = _bytes(lib, kernel_table(lib))
    for k, h in enumerate(hists):
        print(f"# {label} step {k}: " + " ".join(
            f"[itr {x['itr']} ls {x['ls_itr']} iNorm {x['iNorm']:.3e} fNorm {x['fNorm']:.3e}"
            + (f" fallback@{x['fallback_itr']}" if x.get('fallback_itr', -1) >= 0 else "") + "]" for x in h),
            file=sys.stderr, flush=True)
    n_newton = sum(len(h) for h in hists)
    ok = bool(all(x["suc"] for h in hists for x in h))
    return {"elapsed": t1 - t0, "s_per_newton": (t1 - t0) / n_newton, "n_newton": n_newton, "hists": hists,
            "kernels": ktab, "kernels_warmup": warm_tab, "Yn": fn.Yn.clone(), "An": fn.An.clone(),
            "all_converged": ok, "steps": steps, "warmup": warmup}


def roofline(ktab, measured=None, variants=None):
    """Roofline of the kernel with the largest total time among those with
    algorithmic byte counts: achieved = bytes per launch / average duration;
    traffic = the counter-measured HBM bytes per launch of the launch variant
    the autotuner picked, where known."""
    cands = [(v["total_ms"], k) for k, v in ktab.items() if k in HBM_KERNELS and v.get("bytes")]
    if not cands:
        return None
    _, k = max(cands)
    v = ktab[k]
    avg_s = v["total_ms"] / v["count"] * 1e-3
    bpl = v["bytes"] / v["count"]
    ach = bpl / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": _traffic(measured, variants, k), "kernel": k,
            "bytes_per_launch": round(bpl),
            "avg_us": round(avg_s * 1e6, 2)}


def arnoldi_step(ktab, basis):
    """Roofline of the whole Arnoldi step (SpMV + Gram-Schmidt dots + finaliser
    + update) from a kernel table in which every launch was timed: the
    algorithmic bytes of those launches over their summed durations, per
    Arnoldi step (= per update launch)."""
    ks = ARNOLDI_KERNELS[basis]
    if not all(k in ktab and ktab[k].get("count") for k in ks):
        return None
    steps = ktab[ks[3]]["count"]
    t = sum(ktab[k]["total_ms"] for k in ks) * 1e-3
    b = sum(ktab[k].get("bytes", 0.0) for k in ks)
    return {"kernels": ks, "steps": steps, "us_per_step": round(t / steps * 1e6, 2),
            "bytes_per_step": round(b / steps), "achieved": round(b / t / 1e9, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(b / t / 1e9 / HBM_PEAK_GBS, 4),
            "us_per_step_at_peak": round(b / steps / (HBM_PEAK_GBS * 1e9) * 1e6, 2)}


def _traffic(measured, variants, k):
    if not measured or not variants or k not in measured:
        return None
    v = variants.get({"spmv_bcsr": "fp64", "spmv32": "fp32"}[k], -1)
    return measured[k].get(v)


def get_network(a, log, path=None, coarse=None):
    """The learned preconditioner: committed weights, or a bounded training run
    (as the two-level smoother when coarse > 0)."""
    from svfsiplus_amd import lpc as L
    path = path or a.lpc_weights or (DEFAULT_SMOOTHER if a.coarse > 0 else DEFAULT_WEIGHTS)
    coarse = a.coarse if coarse is None else coarse
    if os.path.exists(path):
        with np.load(path, allow_pickle=False) as d:
            meta = {k: d[k].item() for k in d.files if k != "params"}
        return L.LearnedPC.load(path), {"source": os.path.relpath(path, ROOT), **meta}
    t0 = time.perf_counter()
    data = L.build_training_set(a.train_draws, seed=11, k_max=8, log=log, coarse=coarse)
    net = L.LearnedPC(seed=0)
    hist = L.train(net, data, steps=a.train_steps, lr=2e-3, batch=8, log=log, log_every=500, smoother=coarse > 0)
    info = {"source": "trained in this run", "draws": a.train_draws, "coarse": coarse,
            "triples": int(sum(r.shape[0] for _, r, _ in data)), "steps": a.train_steps,
            "final_loss": round(float(sum(hist[-100:]) / max(len(hist[-100:]), 1)), 5),
            "train_s": round(time.perf_counter() - t0, 1)}
    del data
    return net, info


def mixed_draws(a, fn, lib, log):
    """The mixed-precision workload: seeded draws of viscosity (log-uniform in
    [0.02, 0.08]) and inflow peak (uniform in [10, 50]) on the same pipe; per
    draw one time step with the plain FP64 GMRES and one with the FP32-SpMV /
    FP32-basis GMRES refined in FP64 (diagonal preconditioner, true-residual
    restarts), each timed; the mixed result's error against the FP64 one."""
    import torch
    from svfsiplus_amd import _lib, workload as W
    rng = np.random.default_rng(a.c3_seed)
    draws = []
    for d in range(a.c3_draws):
        mu = float(np.exp(rng.uniform(np.log(0.02), np.log(0.08))))
        peak = float(rng.uniform(10.0, 50.0))
        fn.set_physics(W.Physics(mu=mu, peak=peak))
        res = {}
        for label, prec in (("fp64", _lib.PRECISION_FP64), ("mixed", _lib.PRECISION_MIXED)):
            fn.precision, fn.lpc, fn.coarse = prec, None, 0
            fn.reset()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            h = [dict(x) for x in fn.time_step()]
            torch.cuda.synchronize()
            res[label] = (time.perf_counter() - t0, h, fn.Yn.clone(), fn.An.clone())
        (t64, h64, Y64, A64), (t32, h32, Y32, A32) = res["fp64"], res["mixed"]

        def rel(x, ref):
            return float(torch.linalg.vector_norm(x - ref) / torch.linalg.vector_norm(ref))
        rec = {"mu": round(mu, 6), "peak": round(peak, 4),
               "s_per_newton": round(t32 / len(h32), 5), "fp64_s_per_newton": round(t64 / len(h64), 5),
               "speedup_vs_fp64": round((t64 / len(h64)) / (t32 / len(h32)), 3),
               "linear_iterations": [x["ls_itr"] for x in h32], "fp64_linear_iterations": [x["ls_itr"] for x in h64],
               "all_converged": bool(all(x["suc"] for x in h32 + h64)),
               "error_vs_fp64": {"velocity": rel(Y32[:, :3], Y64[:, :3]), "pressure": rel(Y32[:, 3], Y64[:, 3]),
                                 "acceleration": rel(A32[:, :3], A64[:, :3]),
                                 "final_rel_residual": [x["fNorm"] / x["iNorm"] for x in h32],
                                 "final_rel_residual_fp64": [x["fNorm"] / x["iNorm"] for x in h64]}}
        log(f"# mixed draw {d}: mu {mu:.4f} peak {peak:.2f}: {rec['s_per_newton']} s/Newton (fp64 "
            f"{rec['fp64_s_per_newton']}), itr {rec['linear_iterations']} vs {rec['fp64_linear_iterations']}, "
            f"err u {rec['error_vs_fp64']['velocity']:.2e} p {rec['error_vs_fp64']['pressure']:.2e}")
        draws.append(rec)
    fn.set_physics(W.Physics())
    return {"draws": draws, "seed": a.c3_seed,
            "s_per_newton_mean": round(float(np.mean([d["s_per_newton"] for d in draws])), 5),
            "all_converged": bool(all(d["all_converged"] for d in draws)),
            "solver": "GMRES(%d) with FP32 block SpMV and FP32 Krylov basis, FP64 Gram-Schmidt/Givens and FP64 "
                      "true-residual restarts (convergence decided on the true residual), diagonal "
                      "preconditioner; one time step x 3 Newton iterations per draw" % a.krylov}


def configure(fn, method, a, net=None):
    """The solver options of a method on a FluidNewton."""
    from svfsiplus_amd import _lib
    fn.precision = _lib.PRECISION_FP64 if method == "plain" else _lib.PRECISION_MIXED
    fn.coarse = a.coarse if method in ("twolevel", "learned") else 0
    fn.coarse_keep = bool(a.coarse_keep) and fn.coarse > 0
    fn.lpc = net if method == "learned" else None
    fn.net_period = a.net_period


def ns_run(a, log, with_cpu=True):
    """The 15k-tet pipe (n = 8, the C1 case) with the fluid equation's default
    linear solver, FSILS NS (ns_solver.cpp:167-499), with fsils_ls_create's NS
    defaults (ls.cpp:51-60: RI relTol 0.4, absTol 1e-10, 10 outer iterations,
    Krylov 100; GM 1e-2 / 2; CG 0.2 / 500); one time step of 3 Newton
    iterations (outlet backflow face included): seconds per Newton step on the
    device (a warm-up step, then a.ns_steps timed) and the RI/GM/CG iteration
    counts per Newton iteration; then the reference's own CPU path on the same
    three systems (its construct_fluid, b_assem_neu_bc and fsils_solve from
    oracle/_ref, 1 thread, the same parameters), each assembled from the
    device's Newton iterates, timed in full."""
    import torch
    import oracle
    from svfsiplus_amd import _lib, workload as W
    from svfsiplus_amd.newton import FluidNewton, ls_create
    pipe = W.make_pipe(8)
    phys = W.Physics()
    import ctypes as C
    ls0 = _lib.Ls()
    _lib.check(_lib.load().svg_ls_create(C.byref(ls0), _lib.LS_NS))
    fn = FluidNewton(pipe, phys, ls0)
    states, recs = [], []
    els = []
    for k in range(1 + a.ns_steps):
        fn.reset()
        fn.eq.itr = 0
        fn.eq.ok = False
        torch.cuda.synchronize()
        t = time.perf_counter()
        rec = []
        while True:
            if k == 0:
                states.append(None)
            done = fn.newton_iteration()
            ls = fn.ls
            rec.append({"RI": ls.RI.itr, "GM": ls.GM.itr, "CG": ls.CG.itr, "Resm": ls.Resm, "Resc": ls.Resc,
                        "iNorm": ls.RI.iNorm, "fNorm": ls.RI.fNorm, "suc": ls.RI.suc})
            log(f"# ns C1 step {k} Newton {len(rec)}: RI/GM/CG {ls.RI.itr}/{ls.GM.itr}/{ls.CG.itr} suc {ls.RI.suc}")
            if k == 0:
                states[-1] = (fn.Ag.cpu().numpy(), fn.Yg.cpu().numpy())
            if done:
                break
        torch.cuda.synchronize()
        els.append(time.perf_counter() - t)
        recs.append(rec)
    el = float(np.mean(els[1:]))
    out = {"workload": f"VMS pipe n=8: {pipe.nEl} tets, {pipe.nNo} nodes, nnzb {pipe.nnz}; 1 time step x "
                       f"{len(recs[-1])} Newton iterations; FSILS NS solver with fsils_ls_create's NS defaults: "
                       f"RI relTol {ls0.RI.relTol:g} / absTol {ls0.RI.absTol:g} / {ls0.RI.mItr} iterations / Krylov "
                       f"{ls0.RI.sD}, GM {ls0.GM.mItr} / {ls0.GM.relTol:g}, CG {ls0.CG.mItr} / {ls0.CG.relTol:g}",
           "s_per_newton": round(el / len(recs[-1]), 6), "newton_records": recs[-1],
           "reached_relTol": bool(all(r["suc"] for rr in recs for r in rr)), "steps": a.ns_steps, "warmup": 1,
           "note": "on this system the NS solver's outer iteration barely lowers the residual and stops on its "
                   "10-iteration limit (suc 0), as the reference's own solve of the same systems does "
                   "(cpu_reference); with the <Linear_solver> defaults of an XML <LS type=NS> (1000 outer "
                   "iterations, NS_GM/NS_CG 1000 / 1e-2) it does not finish a Newton iteration within minutes"}
    log(f"# ns C1: {out['s_per_newton']} s/Newton, RI/GM/CG {[(r['RI'], r['GM'], r['CG']) for r in recs[-1]]}")
    if with_cpu and oracle.ref_available():
        fi, ge, _ = W.outlet_faces(pipe)
        prm = (ls0.RI.relTol, ls0.RI.absTol, ls0.RI.mItr, ls0.RI.sD, ls0.GM.relTol, ls0.GM.absTol, ls0.GM.mItr,
               ls0.GM.sD, ls0.CG.relTol, ls0.CG.absTol, ls0.CG.mItr)
        t_cpu, crec = 0.0, []
        for Ag, Yg in states:
            t = time.perf_counter()
            R, V = oracle.ref_construct_fluid(pipe, phys, Ag, Yg)
            oracle.ref_b_assem_neu_bc(pipe, phys, Yg, fi, ge, R, V)
            _, _, r = oracle.ref_fsils_solve(pipe, oracle.LS_NS, prm, R, V)
            t_cpu += time.perf_counter() - t
            crec.append((int(r["itr"]), int(r["GM_itr"]), int(r["CG_itr"]), int(r["suc"])))
            log(f"# ns C1 cpu reference Newton {len(crec)}: RI/GM/CG {crec[-1][:3]}")
        cpu = host_cpu()
        out["cpu_reference"] = {"value": round(t_cpu / len(states), 4), "unit": "s/Newton-step", "cores": 1,
                                "kind": "reference", "ri_iterations": [c[0] for c in crec],
                                "gm_iterations": [c[1] for c in crec], "cg_iterations": [c[2] for c in crec],
                                "ri_suc": [c[3] for c in crec], "host_cpu": cpu["model"],
                                "host_mhz": cpu["mhz"],
                                "sample": "the reference's construct_fluid, b_assem_neu_bc and fsils_solve (NS) "
                                          "compiled from its sources, 1 thread, on the three Newton systems of the "
                                          "time step (assembled from the device's iterates), every one timed in full"}
        out["speedup_vs_cpu_reference"] = round(t_cpu / len(states) / out["s_per_newton"], 1)
        log(f"# ns C1 cpu reference: {out['cpu_reference']['value']} s/Newton")
    return out


def c4_run(a, lib, net, log):
    """The 4.2M-tet stenosed pipe (n = 52 by default, stenosis depth 0.5 over z in
    [1.25, 1.75]): one time step of 3 Newton iterations with the twolevel and the
    learned methods (each timed after one warm-up step), then the plain FP64
    GMRES on the first Newton system alone with a budget of c4_plain_restarts
    restart cycles (timed; where it does not reach the tolerance within the
    budget, its time is a lower bound of what it needs)."""
    import torch
    from svfsiplus_amd import _lib, workload as W
    from svfsiplus_amd.newton import FluidNewton, ls_create
    t0 = time.perf_counter()
    pipe = W.make_pipe(a.c4_n, sten_depth=a.c4_sten, sten_z0=1.25, sten_len=0.5)
    fn = FluidNewton(pipe, W.Physics(), ls_create(_lib.LS_GMRES, relTol=a.tol, absTol=1e-30, maxItr=1000,
                                                  dimKry=a.krylov), device_setup=True)
    log(f"# c4 setup {time.perf_counter() - t0:.1f}s: {pipe.nEl} tets, {pipe.nNo} nodes, nnzb {pipe.nnz}")
    out = {"workload": f"stenosed VMS pipe n={a.c4_n} (depth {a.c4_sten:g} over z in [1.25, 1.75]): {pipe.nEl} tets, "
                       f"{pipe.nNo} nodes, nnzb {pipe.nnz}; 1 time step x 3 Newton iterations; GMRES({a.krylov}) "
                       f"rel-tol {a.tol:g}"}
    ac = argparse.Namespace(**vars(a))
    ac.coarse, ac.net_period = a.c4_coarse, a.c4_net_period
    out["coarse"], out["net_period"] = ac.coarse, ac.net_period
    for method in ("twolevel", "learned"):
        configure(fn, method, ac, net)
        els = []
        for _ in range(1 + a.c4_steps):
            fn.reset()
            torch.cuda.synchronize()
            t = time.perf_counter()
            h = [dict(x) for x in fn.time_step()]
            torch.cuda.synchronize()
            els.append(time.perf_counter() - t)
        el = float(np.mean(els[1:]))
        out[method] = {"s_per_newton": round(el / len(h), 5), "linear_iterations": [x["ls_itr"] for x in h],
                       "all_converged": bool(all(x["suc"] for x in h)),
                       "final_rel_residual": [x["fNorm"] / x["iNorm"] for x in h],
                       "fallback_itr": [x.get("fallback_itr", -1) for x in h], "steps": a.c4_steps, "warmup": 1}
        log(f"# c4 {method}: {out[method]['s_per_newton']} s/Newton, itr {out[method]['linear_iterations']}")
    Y_learned = fn.Yn.clone()
    # plain FP64 on the first Newton system within the restart budget
    configure(fn, "plain", a)
    fn.ls = ls_create(_lib.LS_GMRES, relTol=a.tol, absTol=1e-30, maxItr=a.c4_plain_restarts, dimKry=a.krylov)
    fn.reset()
    torch.cuda.synchronize()
    t = time.perf_counter()
    fn.eq.itr = 0
    fn.newton_iteration()
    torch.cuda.synchronize()
    tp = time.perf_counter() - t
    hp = fn.history[-1]
    first = out["learned"]["s_per_newton"]
    out["plain_first_newton"] = {
        "restart_budget": a.c4_plain_restarts, "linear_iterations": hp["ls_itr"], "s": round(tp, 4),
        "converged": bool(hp["suc"]), "rel_residual_reached": hp["fNorm"] / hp["iNorm"],
        "note": "the plain FP64 GMRES on the first Newton system with a budget of %d restart cycles; when it has not "
                "converged, s is a lower bound of its time for that Newton iteration alone" % a.c4_plain_restarts}
    out["speedup_vs_plain_lower_bound"] = round(tp / first, 2) if not hp["suc"] else None
    out["speedup_vs_plain"] = round(tp / first, 2) if hp["suc"] else None
    log(f"# c4 plain (first Newton system, {a.c4_plain_restarts} restarts): {tp:.2f}s, itr {hp['ls_itr']}, "
        f"rel res {hp['fNorm'] / hp['iNorm']:.2e}, converged {bool(hp['suc'])}")
    del Y_learned
    fn = None
    return out


def batch_run(a, log):
    """The batched workload: independent bifurcating-pipe meshes (tet counts
    log-uniform in 5k-80k, daughter angles 25-45 degrees, Murray radii) solved
    together: one union assembly and the batched mixed-precision GMRES (by
    default with the learned preconditioner, per-mesh fallback), each Newton
    iteration as three replayed hipGraphs. Aggregate Newton steps per second."""
    import torch
    from svfsiplus_amd import _lib, batch as B, workload as W
    t0 = time.perf_counter()
    specs = W.bifurcation_batch(a.batch)
    meshes = B.make_batch(specs)
    phys = W.Physics()
    prec = _lib.PRECISION_MIXED if a.batch_precision == "mixed" else _lib.PRECISION_FP64
    net = None
    if a.batch_method == "learned":
        from svfsiplus_amd import lpc as L
        net = L.LearnedPC.load(DEFAULT_WEIGHTS)
    coarse = a.batch_coarse if a.batch_method == "twolevel" else 0
    bn = B.BatchNewton(meshes, phys, sD=a.krylov, rel_tol=a.tol, precision=prec, lpc=net, net_period=a.net_period,
                       coarse=coarse)
    log(f"# batch setup {time.perf_counter() - t0:.1f}s: {len(meshes)} meshes, {bn.pipe.nEl} tets, "
        f"{bn.pipe.nNo} nodes, nnzb {bn.pipe.nnz}")
    for w in range(a.batch_warmup):
        bn.reset()
        bn.time_step(3)
    torch.cuda.synchronize()
    r0 = bn.replays()
    t1 = time.perf_counter()
    recs = []
    for _ in range(a.batch_steps):
        bn.reset()
        recs.append(bn.time_step(3))
    torch.cuda.synchronize()
    el = time.perf_counter() - t1
    syncs = (bn.replays() - r0) / a.batch_steps
    newton = len(meshes) * 3 * a.batch_steps
    itr = [[r["itr"] for r in it] for it in recs[-1]]
    ok = all(r["suc"] for it in recs[-1] for r in it)
    fallbacks = [sum(1 for r in it if r.get("fallback_itr", -1) >= 0) for it in recs[-1]]
    log(f"# batch: {newton} Newton steps in {el:.2f}s; max linear iterations per Newton iteration "
        f"{[max(x) for x in itr]}, all converged {ok}")
    ok = ok and all(r["suc"] for rr in recs for it in rr for r in it)
    out = {"meshes": len(meshes), "tets": int(bn.pipe.nEl), "nodes": int(bn.pipe.nNo), "nnzb": int(bn.pipe.nnz),
           "newton_steps_per_s": round(newton / el, 2), "s_per_time_step": round(el / a.batch_steps, 4),
           "steps": a.batch_steps, "warmup": a.batch_warmup,
           "precision": a.batch_precision, "method": a.batch_method, "all_converged": ok,
           "linear_iterations": {"max": [max(x) for x in itr], "mean": [round(float(np.mean(x)), 1) for x in itr]},
           "restart_cycles_max": [int(np.ceil(max(x) / a.krylov)) for x in itr],
           "host_syncs_per_time_step": syncs,
           "solver": "batched %sGMRES(%d), per-mesh Krylov spaces and convergence%s; each Newton iteration is "
                     "replayed hipGraphs (pici + assembly + preconditioner; the restart cycles, four per replay while "
                     "more than an eighth of the meshes iterate, then one; the corrector), the host reading the "
                     "meshes' states once per replay; block SpMV variant from the autotuner" % (
                         "flexible " if (net or coarse) else "", a.krylov,
                         (", two-level preconditioner (aggregates of about %d nodes, one dense coarse block per mesh "
                          "inverted inside the captured graph; block-Jacobi smoother; per-mesh fallback to the "
                          "diagonal preconditioner)" % coarse) if coarse else "")}
    out["spmv_variants"] = spmv_variants(_lib.load(), bn.fn.lhs)
    if net is not None:
        out["learned_pc"] = {"weights": os.path.relpath(DEFAULT_WEIGHTS, ROOT), "net_period": a.net_period,
                             "meshes_fallen_back": fallbacks}
    if coarse:
        out["twolevel"] = {"coarse": coarse, "meshes_fallen_back": fallbacks}
    u = bn.pipe
    bn.free()
    # the union's mesh setup (block pattern, colouring, colour order, slot map): host and device builds
    from svfsiplus_amd.newton import DeviceLhs, DeviceMesh, device_arrays
    torch.cuda.synchronize()
    t = time.perf_counter()
    W.pattern(u.nNo, u.IEN)
    Lh = DeviceLhs(u)
    Mh = DeviceMesh(Lh, u)
    torch.cuda.synchronize()
    th = time.perf_counter() - t
    t = time.perf_counter()
    d_ien, d_x = device_arrays(u)
    Ld = DeviceLhs(u, d_ien=d_ien)
    Md = DeviceMesh(Ld, u, d_ien=d_ien, d_x=d_x)
    torch.cuda.synchronize()
    td = time.perf_counter() - t
    out["setup_s"] = {"host": round(th, 3), "device": round(td, 3),
                      "colours": {"host": Mh.nColors, "device": Md.nColors},
                      "note": "block pattern + colouring + colour order + slot map of the union (host: C++ on one "
                              "core, then upload; device: upload of IEN and x, then svg_lhs_create_dev and "
                              "svg_msh_create_dev)"}
    log(f"# batch union setup: host {th:.3f}s, device {td:.3f}s")
    for o in (Md, Ld, Mh, Lh):
        o.free()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=34, help="pipe cross-section cells (even; 34 -> 1.2M tets, 52 -> 4.2M)")
    ap.add_argument("--stenosis", type=float, default=0.0,
                    help="sinusoidal stenosis depth (fraction of the radius) over z in [1.25, 1.75]")
    ap.add_argument("--krylov", type=int, default=50)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--method", choices=["compare", "plain", "mixed", "twolevel", "learned"], default="compare")
    ap.add_argument("--lpc-weights", default=None,
                    help="network weights (default: the two-level smoother's, or the stand-alone network's when "
                         "--coarse 0)")
    ap.add_argument("--coarse", type=int, default=64,
                    help="two-level preconditioner of twolevel/learned: node aggregates of about this many nodes "
                         "(0: learned is the stand-alone network after the diagonal scaling)")
    ap.add_argument("--coarse-keep", type=int, default=1,
                    help="1: Newton iterations 2.. of a time step keep the coarse inverse of the first "
                         "(A P and block Jacobi rebuilt every iteration); 0: rebuild it every iteration")
    ap.add_argument("--net-period", type=int, default=8,
                    help="learned: the network preconditions every net-period-th Arnoldi step, the others the "
                         "two-level preconditioner with block Jacobi (or, with --coarse 0, the diagonal scaling)")
    ap.add_argument("--c4", type=int, default=1, help="run the stenosed 4.2M-tet leg (0: skip it)")
    ap.add_argument("--ns", type=int, default=1, help="run the NS-solver leg on the 15k-tet pipe (0: skip it)")
    ap.add_argument("--ns-steps", type=int, default=1)
    ap.add_argument("--c4-n", type=int, default=52)
    ap.add_argument("--c4-sten", type=float, default=0.5)
    ap.add_argument("--c4-steps", type=int, default=1)
    ap.add_argument("--c4-plain-restarts", type=int, default=300)
    ap.add_argument("--c4-coarse", type=int, default=128,
                    help="stenosed leg: nodes per aggregate (733k nodes: 64 would exceed the dense coarse limit)")
    ap.add_argument("--c4-net-period", type=int, default=5)
    ap.add_argument("--train-draws", type=int, default=40)
    ap.add_argument("--train-steps", type=int, default=1500)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--plain-steps", type=int, default=2, help="timed time steps of the plain FP64 leg")
    ap.add_argument("--plain-warmup", type=int, default=1)
    ap.add_argument("--c3-draws", type=int, default=3,
                    help="mixed-precision leg: seeded (mu, inflow peak) draws on the same pipe (0: skip it)")
    ap.add_argument("--c3-seed", type=int, default=20261203)
    ap.add_argument("--batch", type=int, default=256, help="meshes of the batched workload (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=5, help="timed time steps of the batched workload")
    ap.add_argument("--batch-warmup", type=int, default=1)
    ap.add_argument("--batch-precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--batch-method", choices=["plain", "learned", "twolevel"], default="twolevel",
                    help="batched solve: diagonal preconditioner, the learned network alone (flexible GMRES), "
                         "or the two-level preconditioner (one coarse block per mesh)")
    ap.add_argument("--batch-coarse", type=int, default=64, help="twolevel batch: nodes per aggregate")
    ap.add_argument("--no-prof", action="store_true", help="time without per-kernel events")
    ap.add_argument("--kernel-only", choices=["spmv", "spmv32", "lpc", "assembly"], default=None,
                    help="launch only the block SpMV on the assembled system (for counter runs), "
                         "plus a device copy of the same byte count as the calibration kernel")
    ap.add_argument("--launches", type=int, default=20)
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched workload and print its record (a measurement aid; the "
                         "headline line comes from the default run)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (the multi-GPU sharding is a design, not built); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from svfsiplus_amd import _lib, workload as W
    from svfsiplus_amd.newton import FluidNewton, ls_create

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    lib = _lib.load()
    if a.batch_only:
        print(json.dumps({"batch": batch_run(a, log)}))
        return
    t_setup = time.perf_counter()
    pipe = (W.make_pipe(a.n, sten_depth=a.stenosis, sten_z0=1.25, sten_len=0.5) if a.stenosis > 0.0
            else W.make_pipe(a.n))
    phys = W.Physics()
    ls = ls_create(_lib.LS_GMRES, relTol=a.tol, absTol=1e-30, maxItr=1000, dimKry=a.krylov)
    fn = FluidNewton(pipe, phys, ls)
    t_setup = time.perf_counter() - t_setup
    log(f"# setup {t_setup:.1f}s: nNo={pipe.nNo} nEl={pipe.nEl} nnzb={pipe.nnz} colours={fn.msh.nColors}")

    if a.kernel_only == "lpc":
        kernel_only_lpc(a, fn, pipe, lib)
        return
    if a.kernel_only == "assembly":
        kernel_only_assembly(a, fn, pipe, lib)
        return
    if a.kernel_only:
        kernel_only(a, fn, pipe, lib)
        return

    runs = {}
    net, net_info = None, None
    if a.method in ("compare", "plain"):
        configure(fn, "plain", a)
        ps, pw = (a.steps, a.warmup) if a.method == "plain" else (a.plain_steps, a.plain_warmup)
        runs["plain"] = timed_run(fn, lib, ps, pw, not a.no_prof, "plain")
    if a.method == "mixed":
        configure(fn, "mixed", a)
        runs["mixed"] = timed_run(fn, lib, a.steps, a.warmup, not a.no_prof, "mixed")
    if a.method in ("compare", "twolevel") and a.coarse > 0:
        configure(fn, "twolevel", a)
        runs["twolevel"] = timed_run(fn, lib, a.steps, a.warmup, not a.no_prof, "twolevel")
    if a.method in ("compare", "learned"):
        net, net_info = get_network(a, log)
        configure(fn, "learned", a, net)
        runs["learned"] = timed_run(fn, lib, a.steps, a.warmup, not a.no_prof, "learned")
    head = "learned" if "learned" in runs else next(iter(runs))
    R = runs[head]

    traffic = MEASURED_TRAFFIC_N34 if (a.n == 34 and a.stenosis == 0.0) else None
    cpu = None
    if not a.no_cpu:
        log("# cpu baseline ...")
        base = runs.get("plain", R)
        cpu = cpu_baseline(pipe, phys, [x for h in base["hists"] for x in h])

    tl_text = ("two-level preconditioner (pressure aggregation coarse level, aggregates of about %d nodes, dense "
               "Galerkin coarse matrix inverted on the device%s; block-Jacobi smoother)"
               % (a.coarse, " in the first Newton iteration of the time step and kept for the others"
                  if a.coarse_keep else " every Newton iteration"))
    solver = {"plain": "FSILS GMRES (gmres_v), diagonal preconditioner, FP64",
              "mixed": "GMRES with FP32 block SpMV, FP64 Gram-Schmidt/Givens, FP64 true-residual restarts, "
                       "diagonal preconditioner",
              "twolevel": "flexible GMRES with FP32 block SpMV, FP64 Gram-Schmidt/Givens and true-residual "
                          "restarts, " + tl_text,
              "learned": ("flexible GMRES with FP32 block SpMV, FP64 Gram-Schmidt/Givens and true-residual restarts, "
                          + tl_text + ", the learned GNN (3 layers, width 64, BF16/MFMA) as the smoother on every "
                          "%d-th Arnoldi step; fallback after 20 non-decreasing iterations to the block-Jacobi "
                          "smoother, then to the diagonal preconditioner" % a.net_period) if a.coarse > 0 else
                         ("flexible GMRES, learned GNN preconditioner (3 layers, width 64, BF16/MFMA) on every "
                          "%d-th Arnoldi step after the diagonal scaling, FP32 block SpMV, FP64 Gram-Schmidt/Givens "
                          "and true-residual restarts, fallback to the diagonal preconditioner after 20 "
                          "non-decreasing iterations" % a.net_period)}
    out = {
        "metric": "Newton-step s at FP64 rel-res 1e-6, 1.2M-tet VMS pipe: learned-PC vs plain",
        "value": round(R["s_per_newton"], 6),
        "unit": "s/Newton-step",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * R["elapsed"] / a.steps, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "plain" else "f64+f32+bf16",
        "data": "synthetic (seeded pipe generator)",
        "all_converged": R["all_converged"],
        "config": {"workload": f"VMS pipe n={a.n}{f' (stenosis {a.stenosis:g})' if a.stenosis > 0 else ''}: "
                               f"{pipe.nEl} tets, {pipe.nNo} nodes, nnzb {pipe.nnz}; "
                               f"1 time step x 3 Newton iterations; GMRES({a.krylov}) rel-tol {a.tol:g} (FP64 "
                               f"residual of the scaled system)",
                   "method": head, "solver": solver[head], "newton_records": R["hists"][-1],
                   "spmv_variants": spmv_variants(lib, fn.lhs)},
        "roofline": roofline(R["kernels"], traffic, spmv_variants(lib, fn.lhs)),
        "arnoldi_step": arnoldi_step(R["kernels_warmup"] or R["kernels"], "fp64" if head == "plain" else "fp32"),
        "cpu_baseline": cpu,
        "kernels": R["kernels_warmup"] or R["kernels"],
        "kernels_note": "per-kernel event table of the untimed warmup step; the timed steps bracket one launch "
                        "in %d of %s with events (the roofline's durations)" % (PROF_EVERY, ", ".join(ROOFLINE_KERNELS)),
    }
    if "plain" in runs and head != "plain":
        P = runs["plain"]
        out["plain"] = {"value": round(P["s_per_newton"], 6), "newton_records": P["hists"][-1],
                        "steps": P["steps"], "warmup": P["warmup"], "all_converged": P["all_converged"],
                        "roofline": roofline(P["kernels"], traffic, spmv_variants(lib, fn.lhs)),
                        "arnoldi_step": arnoldi_step(P["kernels_warmup"] or P["kernels"], "fp64")}
        out["speedup_vs_plain"] = round(P["s_per_newton"] / R["s_per_newton"], 3)
        # the time step's result against the FP64 plain path's (velocity, pressure, acceleration)
        def rel(x, ref):
            return float(torch.linalg.vector_norm(x - ref) / torch.linalg.vector_norm(ref))
        out["error_vs_fp64"] = {"velocity": rel(R["Yn"][:, :3], P["Yn"][:, :3]),
                                "pressure": rel(R["Yn"][:, 3], P["Yn"][:, 3]),
                                "acceleration": rel(R["An"][:, :3], P["An"][:, :3]),
                                "final_fNorm": R["hists"][-1][-1]["fNorm"],
                                "final_fNorm_fp64": P["hists"][-1][-1]["fNorm"]}
    if "twolevel" in runs and head != "twolevel":
        T = runs["twolevel"]
        out["twolevel"] = {"value": round(T["s_per_newton"], 6), "newton_records": T["hists"][-1],
                           "steps": T["steps"], "warmup": T["warmup"], "all_converged": T["all_converged"],
                           "solver": solver["twolevel"]}
        # what the network gains over the same method without it
        it_t = sum(x["ls_itr"] for x in T["hists"][-1])
        it_l = sum(x["ls_itr"] for x in R["hists"][-1])
        out["network_gain"] = {"time": round(T["s_per_newton"] / R["s_per_newton"], 3),
                               "linear_iterations": round(it_t / max(it_l, 1), 3)}
    if net_info:
        out["learned_pc"] = {**net_info, "net_period": a.net_period, "coarse": a.coarse}
    if a.c4 and a.method == "compare":
        out["c4"] = c4_run(a, lib, net, log)
    if a.c3_draws > 0 and a.method == "compare":
        out["mixed"] = mixed_draws(a, fn, lib, log)
    if a.ns and a.method == "compare":
        out["ns_c1"] = ns_run(a, log, with_cpu=not a.no_cpu)
    if a.batch > 0 and a.method == "compare":
        out["batch"] = batch_run(a, log)
    print(json.dumps(out), flush=True)
    if not R["all_converged"]:
        log("# the headline run's linear solves did not all converge: the value is not a valid measurement")
        sys.exit(3)


if __name__ == "__main__":
    main()
