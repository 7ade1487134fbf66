"""Benchmark: D3Q19 MLUPS and HBM-roofline fraction on one MI355X.

Workload (one step = one fused stream-collide launch): the 224^3 V60 brew
domain in FP32 -- 30 deg cone, outlet hole radius 8, 2-cell porous filter
(Darcy K 1e-3, Forchheimer 0.55), Guo gravity 1e-5, Smagorinsky 0.17,
inlet jet radius 4 at u_z = -0.03.  Populations start at rest and are
resident in HBM before timing.  `value` = fluid-cell updates per second
(MLUPS); solid cells (67 %) are neither read nor written.

Secondary lines in the same JSON object: the all-fluid periodic 224^3 box
(every cell updated), FP16/BF16 storage speed-up over FP32, adjoint cost per
forward step, and the CPU baseline (the oracle's C restatement on the host
cores, bounded sample).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, published
BYTES_PER_UPDATE = {"f32": 152, "f16": 76, "bf16": 76, "f64": 304}
REFERENCE_MLUPS_224 = 159.3    # reference's published 224^3 figure (Apple M-series, Metal)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_steps(ctx, steps, warmup):
    """Warm up, then time `steps` launches on the context's stream with HIP events."""
    import torch

    ctx.step(warmup)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    ctx.step(steps)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return start.elapsed_time(stop) / 1e3, wall


def bench_v60(precision, steps, warmup):
    import torch

    from pourover_amd import workloads as WL
    from pourover_amd.engine import HipContext

    wl = WL.v60_config(224, precision)
    counts = wl.geometry.counts()
    ctx = HipContext(wl.config, wl.geometry.flags)
    ctx.init_equilibrium()
    ev_s, wall = time_steps(ctx, steps, warmup)
    red = ctx.macroscopic(macro=False, reduce=True)
    active = counts["fluid"] + counts["inlet"] + counts["outlet"]
    del ctx
    torch.cuda.empty_cache()
    return {"event_s": ev_s, "wall_s": wall, "active_cells": active, "cells": 224**3, "red": red.tolist(),
            "counts": counts}


def bench_periodic(precision, steps, warmup, n=224):
    import torch

    from pourover_amd import workloads as WL
    from pourover_amd.engine import HipContext

    cfg = WL.taylor_green_config(n, precision)
    rho0, u0 = WL.taylor_green_fields(n)
    ctx = HipContext(cfg)
    ctx.init_equilibrium(torch.from_numpy(rho0), torch.from_numpy(u0))
    ev_s, wall = time_steps(ctx, steps, warmup)
    del ctx
    torch.cuda.empty_cache()
    return {"event_s": ev_s, "wall_s": wall, "cells": n**3}


def copy_calibration(nbytes=1 << 30, reps=5):
    """torch device-to-device copy of `nbytes` (read nbytes + write nbytes per launch):
    the reference point for rocprof FETCH_SIZE / WRITE_SIZE and for the
    practical bandwidth ceiling."""
    import torch

    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        dst.copy_(src)
    stop.record()
    torch.cuda.synchronize()
    s = start.elapsed_time(stop) / 1e3 / reps
    del src, dst
    torch.cuda.empty_cache()
    return {"bytes_per_copy": nbytes, "copy_ms": round(s * 1e3, 4), "copy_GBs": round(2 * nbytes / s / 1e9, 1)}


def cpu_baseline(budget_s=12.0):
    """The oracle's C restatement (OpenMP, all host threads) on the same V60 224^3
    workload for a bounded number of steps."""
    import numpy as np

    from oracle import oracle as O
    from pourover_amd import workloads as WL

    wl = WL.v60_config(224, "f32")
    d = O.desc_from_dict(wl.config.oracle_dict(True))
    f = O.init_equilibrium(224, 224, 224, prec="f32")
    flags = wl.geometry.flags
    counts = wl.geometry.counts()
    active = counts["fluid"] + counts["inlet"] + counts["outlet"]
    O.step(d, f, flags, prec="f32")  # warm (page-in)
    n, t0 = 0, time.perf_counter()
    while True:
        f, _, _, _ = O.step(d, f, flags, prec="f32")
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 50:
            break
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": active * n / el / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": "%d steps of the 224^3 V60 FP32 workload (oracle_step_f32, OpenMP %d threads)" % (n, threads)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--quick", action="store_true", help="headline only (no secondary lines, no CPU baseline)")
    ap.add_argument("--precision", default="f32")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 1 GiB device-to-device copy (known bytes) for counter calibration")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the multi-GPU (z-slab halo exchange) path is not part of this build",
              file=sys.stderr)
        sys.exit(2)

    import __graft_entry__

    __graft_entry__.load_package()
    import torch

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    prec = args.precision
    if args.calibrate_copy:
        out = copy_calibration()
        log("bench: copy calibration %s" % json.dumps(out))
    log("bench: v60 224^3 %s, %d warmup + %d timed steps" % (prec, args.warmup, args.steps))
    r = bench_v60(prec, args.steps, args.warmup)
    k_s = r["event_s"] / args.steps
    mlups = r["active_cells"] / k_s / 1e6
    algo_bytes = BYTES_PER_UPDATE[prec] * r["active_cells"]
    achieved = algo_bytes / k_s / 1e9
    out = {
        "metric": "D3Q19 MLUPS & % HBM roofline at 224^3 (V60 brew domain, fluid-cell updates)",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["wall_s"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(mlups / REFERENCE_MLUPS_224, 2),
        "dtype": prec,
        "data": "synthetic (V60 geometry generated in-tree, fluid at rest + inlet jet)",
        "config": {"workload": "v60_224_%s" % prec, "grid": [224, 224, 224], "active_cells": r["active_cells"],
                   "solid_cells": r["counts"]["solid"], "porous_cells": r["counts"]["porous"], "tau": 0.6,
                   "les_cs": 0.17, "gravity": 1e-5, "darcy_K": 1e-3, "forch_CF": 0.55, "inlet_u": -0.03,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_launch": algo_bytes, "kernel_ms": round(k_s * 1e3, 4)},
        "all_site_mlups": round(r["cells"] / k_s / 1e6, 1),
    }
    if not args.quick:
        log("bench: periodic 224^3 box f32")
        p = bench_periodic("f32", args.steps, args.warmup)
        ps = p["event_s"] / args.steps
        out["periodic_224_f32"] = {"mlups": round(p["cells"] / ps / 1e6, 1), "kernel_ms": round(ps * 1e3, 4),
                                   "achieved_GBs": round(152 * p["cells"] / ps / 1e9, 1),
                                   "frac": round(152 * p["cells"] / ps / 1e9 / HBM_PEAK_GBS, 4)}
        for p16 in ("f16", "bf16"):
            log("bench: v60 224^3 %s" % p16)
            r16 = bench_v60(p16, args.steps, args.warmup)
            s16 = r16["event_s"] / args.steps
            out["%s_storage" % p16] = {"mlups": round(r16["active_cells"] / s16 / 1e6, 1),
                                       "kernel_ms": round(s16 * 1e3, 4), "speedup_vs_f32": round(k_s / s16, 3),
                                       "achieved_GBs": round(76 * r16["active_cells"] / s16 / 1e9, 1)}
        log("bench: cpu baseline")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
