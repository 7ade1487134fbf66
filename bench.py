"""Benchmark: D3Q19 MLUPS and HBM-roofline fraction on one MI355X.

Workload (one step = one fused stream-collide launch): the 224^3 V60 brew
domain in FP32 -- 30 deg cone, outlet hole radius 8, 2-cell porous filter
(Darcy K 1e-3, Forchheimer 0.55), Guo gravity 1e-5, Smagorinsky 0.17,
inlet jet radius 4 at u_z = -0.03.  Populations start at rest and are
resident in HBM before timing, in the compact row-run layout (each row
stores only the run between its first and last non-solid cell).  `value` =
fluid-cell updates per second (MLUPS); the 67 % of cells that are solid get
no threads except the solid cells inside a row's run.

Secondary lines in the same JSON object: the all-fluid periodic 224^3 box
(every cell updated), FP16/BF16 storage speed-up over FP32, the two-way
coupled step with 50 000 particles, adjoint cost per forward step, field
diagnostics per call, the thermal D3Q7 step, the phase-field passes, the
body-force producers, the whole main-loop step, and the CPU baseline (C
restatements on the host cores, bounded samples).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick] [--precision P]
       [--coupled-only | --adjoint-only | --diagnostics-only | --thermal-only | --multiphase-only |
        --forcing-only | --main-loop-only | --periodic-only]
       [--calibrate-copy]
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
# the reference's only published 224^3 figure (Apple M-series, Metal); its own
# numbers contradict each other, so it is quoted beside the result, not divided by
REFERENCE_MLUPS_224 = 159.3


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
    layout = "compact row runs" if ctx.compact else "dense"
    so_s = time_stream_only(ctx, steps, warmup)
    del ctx
    torch.cuda.empty_cache()
    return {"event_s": ev_s, "wall_s": wall, "active_cells": active, "cells": 224**3, "red": red.tolist(),
            "counts": counts, "layout": layout, "stream_only_s": so_s}


def time_stream_only(ctx, steps, warmup):
    """Seconds per collisionless streaming pass (pol_stream_only) of the
    same state and layout: the step's loads and stores without the collision,
    the access pattern's bandwidth ceiling for the step kernel."""
    import torch

    ctx.stream_only(warmup)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    ctx.stream_only(steps)
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / 1e3 / steps


def pattern_ceiling(step_s, so_s, algo_bytes):
    """The collisionless pass as the step's access-pattern ceiling."""
    return {"kernel": "collisionless streaming pass (pol_stream_only), same layout and launch",
            "kernel_ms": round(so_s * 1e3, 4), "achieved_GBs": round(algo_bytes / so_s / 1e9, 1),
            "step_frac_of_ceiling": round(so_s / step_s, 3)}


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


def bench_coupled(precision, steps, warmup, count=50_000):
    """Two-way coupled step of the 224^3 cup with the 50 000-particle seed:
    particle coupling, fused stream-collide (writing rho/u for the next
    interpolation) and particle advance, all in one native loop."""
    import torch

    from pourover_amd import workloads as WL
    from pourover_amd.engine import HipContext
    from pourover_amd.particles import HipCoffeeParticleSystem, particle_gravity

    wl = WL.v60_config(224, precision, particle_field=True)
    seedp = WL.seed_particles(wl.geometry, count)
    ctx = HipContext(wl.config, wl.geometry.flags)
    ctx.init_equilibrium()
    ps = HipCoffeeParticleSystem(ctx, seedp.pos, seedp.radius, seedp.mass, gravity=particle_gravity(1e-5, 1.2))
    ps.prime()
    ps.coupled_step(nsteps=warmup)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    ps.coupled_step(nsteps=steps)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / steps
    counts = wl.geometry.counts()
    active = counts["fluid"] + counts["inlet"] + counts["outlet"]
    n_active = int(ps.active.sum().item())
    del ps, ctx
    torch.cuda.empty_cache()
    return {"ms_per_step": round(ms, 4), "mlups": round(active / ms / 1e3, 1), "particles": count,
            "active_particles_after": n_active, "precision": precision}


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


def _omp_threads(n):
    """Set the OpenMP team size of the oracle library (shared libgomp runtime)."""
    import ctypes

    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))
        return True
    except OSError:
        return False


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform string."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or platform.machine()


def cpu_baseline(budget_s=10.0, single_budget_s=6.0):
    """The C/OpenMP restatement of the reference CPU backend
    (cpu_backend.py:79-237: BGK collide + pull stream, FP32, SoA) on the
    224^3 V60 solid mask, a bounded number of steps, on the host cores; plus
    a one-thread figure (the reference calls its CPU backend a single-thread
    reference, cpu_backend.py:36)."""
    import numpy as np

    from oracle import oracle as O
    from pourover_amd import workloads as WL

    wl = WL.v60_config(224, "f32")
    counts = wl.geometry.counts()
    active = counts["fluid"] + counts["inlet"] + counts["outlet"]
    solid = (wl.geometry.kind == 1).astype(np.uint8)
    n3 = (224, 224, 224)
    f = np.zeros((19,) + n3, dtype=np.float32)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float32)
    f[:] = w[:, None, None, None]
    f_new = np.empty_like(f)
    rho = np.empty(n3, dtype=np.float32)
    u = np.empty((3,) + n3, dtype=np.float32)

    def run(budget, max_steps):
        O.cpu_backend_step(f, f_new, rho, u, solid, 0.6, prec="f32")  # warm (page-in)
        n, t0 = 0, time.perf_counter()
        while True:
            O.cpu_backend_step(f, f_new, rho, u, solid, 0.6, prec="f32")
            n += 1
            el = time.perf_counter() - t0
            if el > budget or n >= max_steps:
                return n, el

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    _omp_threads(threads)
    n, el = run(budget_s, 40)
    out = {"value": round(active * n / el / 1e6, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
           "cpu_model": cpu_model(),
           "sample": "%d steps of the 224^3 V60 mask, FP32, C/OpenMP restatement of the reference CPU backend "
                     "(collide + pull stream), %d threads; MLUPS over the %d non-solid cells" % (n, threads, active),
           "all_site_mlups": round(224**3 * n / el / 1e6, 2)}
    if _omp_threads(1):
        n1, el1 = run(single_budget_s, 4)
        out["single_thread"] = {"value": round(active * n1 / el1 / 1e6, 3), "cores": 1, "steps": n1}
        _omp_threads(threads)
    # the same V60 physics as the GPU step (oracle_step_f32: walls, porous
    # layer, LES, Guo gravity, inlet/outlet), same threads
    d = O.desc_from_dict(wl.config.oracle_dict(True))
    g = O.init_equilibrium(224, 224, 224, prec="f32")
    flags = wl.geometry.flags
    g, _, _, _ = O.step(d, g, flags, prec="f32")
    n2, t0 = 0, time.perf_counter()
    while True:
        g, _, _, _ = O.step(d, g, flags, prec="f32")
        n2 += 1
        el2 = time.perf_counter() - t0
        if el2 > budget_s or n2 >= 40:
            break
    out["full_physics_port"] = {"value": round(active * n2 / el2 / 1e6, 2), "cores": threads, "steps": n2,
                                "sample": "oracle_step_f32 on the V60 workload (the GPU step's physics)"}
    return out


def bench_adjoint(n=128, steps=100, reps=200):
    """Adjoint cost at the differentiation workload's grid (scaled V60, FP32):
    kernel time of one reverse step over one forward step, both from HIP
    events on the launch stream -- with every forward state kept (the
    default store at this size) the adjoint kernel alone, with checkpoints
    every 50 steps one forward recompute more -- and the whole gradient over
    a plain forward run (wall), for both stores."""
    import torch

    from pourover_amd import _lib as L
    from pourover_amd import workloads as WL
    from pourover_amd.adjoint import AdjointProblem

    wl = WL.v60_config(n, "f32", perm_field=True)
    K = WL.permeability_field(wl.geometry)
    prob = AdjointProblem(wl, K, steps=steps, window=20, checkpoint=50)
    ctx, lib = prob.ctx, prob.ctx.lib
    ctx.init_equilibrium()
    ctx.step(600)  # warm-up: the engine clock rises over the first tens of milliseconds
    ev = lambda: torch.cuda.Event(enable_timing=True)
    a, b, c, d = ev(), ev(), ev(), ev()
    torch.cuda.synchronize()
    a.record()
    ctx.step(reps)
    b.record()
    state = prob.seg[0]
    state.copy_(ctx.f_current)
    lib.pol_adjoint_reset(ctx.handle, L.stream_handle())
    for _ in range(50):  # warm-up: the first launch also loads the kernel's code object
        L.check(lib.pol_adjoint_step(ctx.handle, L.ptr(state), 1, L.stream_handle()), ctx.handle, "pol_adjoint_step")
    torch.cuda.synchronize()
    c.record()
    for _ in range(reps):
        st = lib.pol_adjoint_step(ctx.handle, L.ptr(state), 1, L.stream_handle())
        L.check(st, ctx.handle, "pol_adjoint_step")
    d.record()
    torch.cuda.synchronize()
    fwd_ms = a.elapsed_time(b) / reps
    adj_ms = c.elapsed_time(d) / reps
    J, _, _ = prob.gradient()
    # The checkpointed reverse sweep of the differentiation workload (500
    # steps, checkpoints every 50) as it runs: per reverse step (segment
    # rebuild + adjoint kernels), from HIP events around the sweep.
    big = AdjointProblem(wl, K, steps=500, window=100, checkpoint=50)
    big.forward(checkpoints=True)
    L.check(lib.pol_adjoint_reset(big.ctx.handle, L.stream_handle()), big.ctx.handle, "pol_adjoint_reset")
    big.reverse_sweep()  # warm-up (kernels and the second stream)
    big.forward(checkpoints=True)
    L.check(lib.pol_adjoint_reset(big.ctx.handle, L.stream_handle()), big.ctx.handle, "pol_adjoint_reset")
    e0, e1 = ev(), ev()
    torch.cuda.synchronize()
    e0.record()
    big.reverse_sweep()
    e1.record()
    torch.cuda.synchronize()
    rev_step_ms = e0.elapsed_time(e1) / 500
    del big
    torch.cuda.empty_cache()
    t_grad = prob.timings["forward_s"] + prob.timings["reverse_s"]
    t_fwd = prob.forward_time()
    del prob
    torch.cuda.empty_cache()
    # every forward state kept (no recomputation in the reverse sweep)
    prob = AdjointProblem(wl, K, steps=steps, window=20, checkpoint=None)
    prob.gradient()
    t_grad_all = prob.timings["forward_s"] + prob.timings["reverse_s"]
    del prob
    torch.cuda.empty_cache()
    out = {"grid": [n, n, n], "precision": "f32", "forward_kernel_ms": round(fwd_ms, 4),
           "adjoint_kernel_ms": round(adj_ms, 4),
           "default_state_store": "every state kept (checkpoint='auto': the states fit in HBM)",
           "adjoint_step_over_forward": round(adj_ms / fwd_ms, 3),
           "reverse_step_ms_checkpointed": round(rev_step_ms, 4),
           "adjoint_step_over_forward_checkpointed": round(rev_step_ms / fwd_ms, 3),
           "adjoint_step_over_forward_checkpointed_serial": round((fwd_ms + adj_ms) / fwd_ms, 3),
           "gradient_over_forward_wall": round(t_grad_all / t_fwd, 2),
           "gradient_over_forward_wall_checkpointed": round(t_grad / t_fwd, 2),
           "steps": steps, "checkpoint_every": 50,
           "note": "with every forward state kept a reverse step is the adjoint kernel alone; with "
                   "checkpoints every 50 steps (500-step sweep, measured) a reverse step also rebuilds one forward "
                   "state; _serial is the sum of one adjoint and one forward kernel"}
    return out


def bench_diagnostics(reps=200, host=True):
    """Field diagnostics at 224^3 (V60, FP32): one pol_field_stats call (the
    statistics every reference diagnostic reads: rho / u / flags pass, FP64
    block partials, 43 doubles to the host) timed per call, against the
    reference's host path -- rho and u copied to the host, then the numpy
    diagnostics (oracle/diagnostics.py: quality, conservation, flow, V60
    physics) -- timed once."""
    import numpy as np
    import torch

    from pourover_amd import diagnostics as D
    from pourover_amd import workloads as WL
    from pourover_amd.engine import HipContext

    wl = WL.v60_config(224, "f32")
    ctx = HipContext(wl.config, wl.geometry.flags)
    ctx.init_equilibrium()
    ctx.step(300)  # also brings the engine clock up to its working value
    ctx.macroscopic(macro=True)
    torch.cuda.synchronize()
    for _ in range(50):
        D.field_statistics(ctx)
    t0 = time.perf_counter()
    for _ in range(reps):
        D.field_statistics(ctx)
    call_ms = (time.perf_counter() - t0) / reps * 1e3
    # device time per call from HIP events on the launch stream: the statistics
    # pass and its fold (plus the 344-byte copy of the results)
    dev_ms = 0.0
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        D.field_statistics(ctx)
        e1.record()
        torch.cuda.synchronize()
        dev_ms += e0.elapsed_time(e1) / reps
    n = wl.config.cells
    algo = 20 * n  # rho 4 B + u 12 B + flags 4 B per cell
    out = {"grid": [224, 224, 224], "ms_per_call": round(call_ms, 4), "algorithmic_bytes": algo,
           "achieved_GBs_wall": round(algo / call_ms / 1e6, 1), "frac_wall": round(algo / call_ms / 1e6 / HBM_PEAK_GBS, 4),
           "device_ms_per_call": round(dev_ms, 4), "achieved_GBs_device": round(algo / dev_ms / 1e6, 1),
           "frac_device": round(algo / dev_ms / 1e6 / HBM_PEAK_GBS, 4)}
    if host:
        sys.path.insert(0, ROOT)
        from oracle import diagnostics as OD

        t0 = time.perf_counter()
        rho = ctx.rho.cpu().numpy().transpose(2, 1, 0)
        u = ctx.u.cpu().numpy().transpose(3, 2, 1, 0)
        t1 = time.perf_counter()
        OD.lbm_quality(rho, u)
        OD.conservation(rho, u, {"initial_mass": 0.0, "accumulated_inflow": 0.0, "prev_momentum": np.zeros(3)}, 0.0)
        OD.flow_analysis(u)
        OD.v60_physics(rho, u, 0.14 / 224, 0.01)
        t2 = time.perf_counter()
        out["host_path_ms"] = round((t2 - t0) * 1e3, 1)
        out["host_copy_ms"] = round((t1 - t0) * 1e3, 1)
        out["host_path_note"] = "rho / u to host + numpy diagnostics (the reference's to_numpy path), one core"
    del ctx
    torch.cuda.empty_cache()
    return out


# thermal D3Q7 algorithmic bytes per cell and step (FP32): collide reads g
# (7 x 4) and the heat source (4) and writes g (7 x 4); the temperature pass
# reads g (7 x 4) and writes T (4).  Convection adds T (4, the stencil counted
# once), u (3 x 4) and the heat-source write (4) to the collide pass (the
# base heat source is read in place of the heat source) and runs the
# temperature pass every step: 112 B.  Pure diffusion stores T in the next
# collide pass (+4 B) and runs the temperature pass once per run: 64 B per
# step + 32 B per run.
def thermal_bytes(conv: bool, steps: int) -> float:
    return 112.0 if conv else 64.0 + 32.0 / steps


def bench_thermal(n=224, steps=200, warmup=20, cpu=True):
    """Thermal D3Q7 step at n^3, FP32 (thermal_lbm.py:396-427 without the
    per-step flux pass): pure diffusion, and with the -u.grad T convection
    source in the coupled loop's form (each step starts from the base heat
    source, thermal_fluid_coupled.py:193-199).  Timed with HIP events on the stream the kernels run on; the
    numpy restatement (oracle/thermal.py, one core) on a 128^3 sample is the
    CPU baseline."""
    import numpy as np
    import torch

    from pourover_amd import thermal as TH

    th = TH.ThermalLBM(grid=(n, n, n), precision="f32")
    th.complete_initialization(25.0, 93.0, n // 4)
    g = torch.Generator(device="cuda").manual_seed(3)
    th._vel.copy_((torch.rand(th._vel.shape, generator=g, device="cuda") - 0.5) * 0.02)
    out = {"grid": [n, n, n], "omega": round(th.omega_thermal, 6), "steps": steps}
    for conv in (False, True):
        th.enable_convection_coupling(conv)
        if conv:  # the coupled loop's form: each step starts from the base heat source
            th.use_base_heat_source(torch.zeros_like(th._hs))
        th.run(warmup)
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        ok = th.run(steps)
        stop.record()
        torch.cuda.synchronize()
        sec = start.elapsed_time(stop) / 1e3 / steps
        algo = int(thermal_bytes(conv, steps) * n ** 3)
        out["convection" if conv else "diffusion"] = {
            "ms_per_step": round(sec * 1e3, 4), "mlups": round(n ** 3 / sec / 1e6, 1),
            "algorithmic_bytes_per_step": algo, "achieved_GBs": round(algo / sec / 1e9, 1),
            "frac": round(algo / sec / 1e9 / HBM_PEAK_GBS, 4), "stable": bool(ok)}
    del th
    torch.cuda.empty_cache()
    if cpu:
        sys.path.insert(0, ROOT)
        from oracle import thermal as OT

        m = 128
        rng = np.random.default_rng(1)
        T = rng.uniform(20, 90, (m, m, m)).astype(np.float32)
        st = {"g": OT.equilibrium(T), "T": T, "heat_source": np.zeros_like(T),
              "u": (rng.random((m, m, m, 3), dtype=np.float32) - 0.5) * 0.02}
        ns, t0 = 0, time.perf_counter()
        while True:
            st = OT.step(st, 1.0 / 0.577, convection=True)
            ns += 1
            el = time.perf_counter() - t0
            if el > 8.0 or ns >= 30:
                break
        out["cpu_baseline"] = {"value": round(m ** 3 * ns / el / 1e6, 2), "unit": "MLUPS", "cores": 1, "kind": "port",
                               "sample": "%d convection steps of a %d^3 box, numpy restatement of thermal_lbm.py, "
                                         "one core" % (ns, m)}
    return out


def bench_multiphase(n=224, reps=100, warmup=10, cpu=True):
    """Phase field on the 224^3 V60 domain, FP32 (multiphase_3d.py:389-421
    in the main loop's order, main.py:769-839): the pre-collision pass
    (gradients, curvature, CSF force, body_force += F / rho) and the phase
    step (advection, M lap(mu), separation, copy, density / phase), timed
    with HIP events on the stream they run on; then the whole main-loop
    step (clear body force, pre-collision pass, stream-collide writing rho /
    u, phase step).  Initial state: water below z = n/4 and a droplet of
    radius n/10, after 20 fluid steps.  CPU baseline: the numpy
    restatement (oracle/multiphase.py, one core) on a 96^3 sample."""
    import numpy as np
    import torch

    from pourover_amd import _lib as L
    from pourover_amd import multiphase as MP
    from pourover_amd import workloads as WL
    from pourover_amd.engine import HipContext

    wl = WL.v60_config(n, "f32", force_field=True)
    ctx = HipContext(wl.config, wl.geometry.flags)
    ctx.init_equilibrium()
    ctx.step(20)
    ctx.macroscopic(macro=True)
    mp = MP.MultiphaseFlow3D(ctx)
    z = torch.arange(n, device="cuda", dtype=torch.float32).view(-1, 1, 1)
    y = torch.arange(n, device="cuda", dtype=torch.float32).view(1, -1, 1)
    x = torch.arange(n, device="cuda", dtype=torch.float32).view(1, 1, -1)
    r = torch.sqrt((x - n / 2) ** 2 + (y - n / 2) ** 2 + (z - 0.6 * n) ** 2)
    phi = torch.maximum(torch.tanh((n / 4 - z) / 2), torch.tanh((n / 10 - r) / 2))
    mp._phi[0].copy_(phi.expand(n, n, n))
    mp._phi_new.fill_(-1.0)
    mp.standardize_initial_state(force_dry_state=False)
    torch.cuda.synchronize()

    def timed(fn, k):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(k):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / 1e3 / k

    def phase_step():
        L.check(mp.lib.pol_multiphase_step(ctx.handle, L.stream_handle()), ctx.handle, "pol_multiphase_step")
        mp._cur ^= 1

    def main_loop_step():
        ctx.force.zero_()
        mp.accumulate_surface_tension_pre_collision()
        ctx.step(1, L.POL_WANT_MACRO)
        phase_step()

    t_if = timed(mp.accumulate_surface_tension_pre_collision, reps)
    t_ph = timed(phase_step, reps)
    t_loop = timed(main_loop_step, reps)
    t_fluid = timed(lambda: ctx.step(1, L.POL_WANT_MACRO), reps)
    N = n ** 3
    solid = torch.from_numpy(wl.geometry.kind == L.POL_KIND_SOLID).to("cuda")
    nsolid = int(solid.sum().item())
    n_if = int(((mp._phi[mp._cur].abs() < 0.9) & ~solid).sum().item())
    # flags everywhere; phi + body force read on fluid cells; rho read and the
    # body force written where a force can arise (fluid cells with |phi| < 0.9)
    b_if = 4 * N + 16 * (N - nsolid) + 16 * n_if
    b_ph = 32 * N                       # phi, mu, u read; phi, rho, phase written
    out = {"grid": [n, n, n], "cells": N, "solid_cells": nsolid, "interface_fluid_cells": n_if,
           "pre_collision_pass": {"ms": round(t_if * 1e3, 4), "algorithmic_bytes": b_if,
                                  "achieved_GBs": round(b_if / t_if / 1e9, 1),
                                  "frac": round(b_if / t_if / 1e9 / HBM_PEAK_GBS, 4)},
           "phase_step": {"ms": round(t_ph * 1e3, 4), "algorithmic_bytes": b_ph,
                          "achieved_GBs": round(b_ph / t_ph / 1e9, 1),
                          "frac": round(b_ph / t_ph / 1e9 / HBM_PEAK_GBS, 4)},
           "main_loop_step_ms": round(t_loop * 1e3, 4),
           "fluid_step_with_macro_ms": round(t_fluid * 1e3, 4),
           "interface": mp.get_interface_statistics()}
    del mp, ctx
    torch.cuda.empty_cache()
    if cpu:
        sys.path.insert(0, ROOT)
        from oracle import multiphase as OM

        m = 96
        g = np.meshgrid(*(np.arange(m, dtype=np.float32),) * 3, indexing="ij")
        rr = np.sqrt((g[0] - m / 2) ** 2 + (g[1] - m / 2) ** 2 + (g[2] - 0.6 * m) ** 2)
        ph = np.maximum(np.tanh((m / 4 - g[2]) / 2), np.tanh((m / 10 - rr) / 2)).astype(np.float32)
        rng = np.random.default_rng(2)
        s = OM.new_state(ph, (rng.random((m, m, m, 3), dtype=np.float32) - 0.5) * 0.02,
                         np.ones((m, m, m), np.float32), np.zeros((m, m, m), np.uint8))
        ns, t0 = 0, time.perf_counter()
        while True:
            OM.accumulate_surface_tension_pre_collision(s)
            OM.step(s, 20, precollision_applied=True)
            ns += 1
            el = time.perf_counter() - t0
            if el > 8.0 or ns >= 30:
                break
        out["cpu_baseline"] = {"value": round(m ** 3 * ns / el / 1e6, 2), "unit": "M cell updates/s", "cores": 1,
                               "kind": "port",
                               "sample": "%d pre-collision passes + steps of a %d^3 box, numpy restatement of "
                                         "multiphase_3d.py, one core" % (ns, m)}
    return out


def bench_forcing(n=224, reps=100, warmup=10):
    """Body-force producers on the 224^3 V60 domain, FP32, one call each:
    the pour force and the pour phase change (nozzle at the cup axis,
    bounding box of the pour region only), the pressure drive in force mode
    (gradient, limit and accumulation in one pass), the density drive, and
    the filter blockage update over the paper cells.  HIP events on the
    stream they run on."""
    import numpy as np
    import torch

    from pourover_amd import _lib as L
    from pourover_amd import forcing as FC
    from pourover_amd import workloads as WL
    from pourover_amd.engine import HipContext

    wl = WL.v60_config(n, "f32", force_field=True)
    ctx = HipContext(wl.config, wl.geometry.flags)
    ctx.init_equilibrium()
    ctx.step(20)
    ctx.macroscopic(macro=True)
    pour = FC.PrecisePouringSystem(ctx)
    pour.start_pouring()
    phi = torch.full((n, n, n), -1.0, device="cuda")
    drive = FC.PressureGradientDrive(ctx)
    drive.activate_force_drive(True)
    filt = FC.FilterPaperSystem(ctx, wl.geometry.porous)
    torch.cuda.synchronize()

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / 1e3 / reps

    def density():
        L.check(ctx.lib.pol_density_drive(ctx.handle, drive._target.data_ptr(), 0.025, 0, 0.001, 0.5, 2.0,
                                          L.stream_handle()), ctx.handle, "pol_density_drive")

    N = n ** 3
    nsolid = int((wl.geometry.kind == L.POL_KIND_SOLID).sum())
    nf = N - nsolid
    nzone = int(np.count_nonzero(wl.geometry.porous))
    out = {"grid": [n, n, n], "fluid_cells": nf, "filter_cells": nzone}
    for name, fn, algo in (
            ("pour_force", lambda: pour.apply_pouring_force(1.0), None),
            ("pour_phase", lambda: pour.apply_gradual_phase_change(phi, None, 1.0), None),
            ("pressure_force_drive", drive.apply_force_drive, 4 * N + 40 * nf),
            ("density_drive", density, 4 * N + 12 * nf),
            ("filter_blockage", filt.update_dynamic_resistance, 20 * nzone)):
        t = timed(fn)
        rec = {"ms": round(t * 1e3, 4)}
        if algo:
            rec.update({"algorithmic_bytes": algo, "achieved_GBs": round(algo / t / 1e9, 1),
                        "frac": round(algo / t / 1e9 / HBM_PEAK_GBS, 4)})
        out[name] = rec
    del ctx
    torch.cuda.empty_cache()
    return out


# The pour fills the cup as the loop runs, so the warm-up fixes the state that
# is timed: after 12 steps about 1.03 M water cells (0.350 ms per step), after
# 100 steps 3.71 M (0.365 ms).  The 12-step state is the one measured.
def bench_main_loop(n=224, reps=50, warmup=12):
    """The reference's whole main-loop step (main.py:769-839, from step 11
    on) on the 224^3 V60 domain, FP32, every part on the device: clear the
    body force, pour force and phase change, pressure drive (force mode)
    with the pre-collision surface tension in the same pass (from step 11
    on), the stream-collide step writing rho / u,
    the filter blockage update and the phase step.  HIP events around
    `reps` consecutive steps."""
    import torch

    from pourover_amd import _lib as L
    from pourover_amd import forcing as FC
    from pourover_amd import multiphase as MP
    from pourover_amd import workloads as WL
    from pourover_amd.engine import HipContext

    wl = WL.v60_config(n, "f32", force_field=True)
    ctx = HipContext(wl.config, wl.geometry.flags)
    ctx.init_equilibrium()
    mp = MP.MultiphaseFlow3D(ctx)
    mp.init_phase_field()           # the dry cup the boundary set-up leaves (boundary_conditions.py:478-481)
    mp.standardize_initial_state(force_dry_state=True)
    pour = FC.PrecisePouringSystem(ctx)
    pour.start_pouring(pattern="spiral")
    drive = FC.PressureGradientDrive(ctx)
    drive.activate_force_drive(True)
    paper = FC.FilterPaperSystem(ctx, wl.geometry.porous)
    count = [0]

    def step():
        count[0] += 1
        ctx.force.zero_()
        pour.apply_pouring_force(1.0)
        pour.apply_gradual_phase_change(mp._phi[mp._cur], None, 1.0)
        if count[0] > 10:           # pressure drive and surface tension in one pass
            drive.apply_with_surface_tension(mp, count[0])
        else:
            drive.apply(count[0])
        ctx.step(1, L.POL_WANT_MACRO)
        paper.update_dynamic_resistance()
        mp.step(count[0], precollision_applied=True)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    a.record()
    for _ in range(reps):
        step()
    b.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps
    ms = a.elapsed_time(b) / reps
    water = float((mp._phi[mp._cur] > 0).sum().item())
    out = {"grid": [n, n, n], "steps": reps, "ms_per_step": round(ms, 4), "wall_ms_per_step": round(wall * 1e3, 4),
           "fluid_cell_mlups": round((wl.geometry.counts()["fluid"] + wl.geometry.counts()["inlet"] +
                                      wl.geometry.counts()["outlet"]) / (ms / 1e3) / 1e6, 1),
           "water_cells_after": int(water),
           "parts": "clear, pour force + phase, pressure drive, surface tension, stream-collide (+rho/u), "
                    "filter blockage, phase step"}
    del mp, ctx
    torch.cuda.empty_cache()
    return out


def counter_traffic(precision):
    """HBM bytes per launch of the stream-collide kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors a
    1 GiB device copy reads for its known bytes), or None."""
    path = os.path.join(ROOT, "profiles", "stream_collide_v60_224_%s_counters.md" % precision)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":", 1)[1]))
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # An idle MI355X starts at a low engine clock and reaches its working
    # clock only after some tens of milliseconds of steps: timed right after
    # 20 warm-up steps, 100 FP16 V60 steps took 0.0663 ms each against
    # 0.0610 ms over 1600 (FP32: 0.1080 against 0.1056 ms).  The defaults warm
    # up for 300 steps (20-30 ms) and time 1000.
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--quick", action="store_true", help="headline only (no secondary lines, no CPU baseline)")
    ap.add_argument("--precision", default="f32")
    ap.add_argument("--coupled-only", action="store_true",
                    help="only the two-way coupled 224^3 step with 50 000 particles (for profiling)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the 128^3 adjoint measurement (for profiling)")
    ap.add_argument("--diagnostics-only", action="store_true",
                    help="only the 224^3 field-diagnostics measurement (for profiling)")
    ap.add_argument("--thermal-only", action="store_true",
                    help="only the 224^3 thermal D3Q7 measurement (for profiling)")
    ap.add_argument("--multiphase-only", action="store_true",
                    help="only the 224^3 phase-field measurement (for profiling)")
    ap.add_argument("--forcing-only", action="store_true",
                    help="only the 224^3 body-force producer measurement (for profiling)")
    ap.add_argument("--periodic-only", action="store_true",
                    help="only the all-fluid periodic 224^3 box (dense layout) measurement (for profiling)")
    ap.add_argument("--main-loop-only", action="store_true",
                    help="only the 224^3 whole main-loop step measurement")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 1 GiB device-to-device copy (known bytes) for counter calibration")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the multi-GPU (z-slab halo exchange) path is not implemented in this package",
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
    if args.coupled_only:
        log("bench: coupled 224^3 %s + 50k particles" % prec)
        print(json.dumps({"coupled_224_50k_particles": bench_coupled(prec, args.steps, args.warmup)}), flush=True)
        return
    if args.adjoint_only:
        log("bench: adjoint 128^3 f32")
        print(json.dumps({"adjoint": bench_adjoint()}), flush=True)
        return
    if args.diagnostics_only:
        log("bench: field diagnostics 224^3 f32")
        print(json.dumps({"diagnostics_224_f32": bench_diagnostics(host=False)}), flush=True)
        return
    if args.multiphase_only:
        log("bench: multiphase 224^3 f32")
        print(json.dumps({"multiphase_224_f32": bench_multiphase(reps=args.steps, warmup=args.warmup, cpu=False)}),
              flush=True)
        return
    if args.periodic_only:
        log("bench: periodic 224^3 box %s" % prec)
        p = bench_periodic(prec, args.steps, args.warmup)
        ps = p["event_s"] / args.steps
        nb = BYTES_PER_UPDATE[prec] * p["cells"]
        print(json.dumps({"periodic_224_%s" % prec: {"mlups": round(p["cells"] / ps / 1e6, 1),
                                                      "kernel_ms": round(ps * 1e3, 4),
                                                      "achieved_GBs": round(nb / ps / 1e9, 1),
                                                      "frac": round(nb / ps / 1e9 / HBM_PEAK_GBS, 4)}}), flush=True)
        return
    if args.main_loop_only:
        log("bench: main-loop step 224^3 f32")
        print(json.dumps({"main_loop_224_f32": bench_main_loop(reps=args.steps, warmup=args.warmup)}), flush=True)
        return
    if args.forcing_only:
        log("bench: body-force producers 224^3 f32")
        print(json.dumps({"forcing_224_f32": bench_forcing(reps=args.steps, warmup=args.warmup)}), flush=True)
        return
    if args.thermal_only:
        log("bench: thermal D3Q7 224^3 f32")
        print(json.dumps({"thermal_224_f32": bench_thermal(steps=args.steps, warmup=args.warmup, cpu=False)}),
              flush=True)
        return
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
        "vs_baseline": None,
        "dtype": prec,
        "data": "synthetic (V60 geometry generated in-tree, fluid at rest + inlet jet)",
        "config": {"workload": "v60_224_%s" % prec, "layout": r["layout"], "grid": [224, 224, 224],
                   "active_cells": r["active_cells"],
                   "solid_cells": r["counts"]["solid"], "porous_cells": r["counts"]["porous"], "tau": 0.6,
                   "les_cs": 0.17, "gravity": 1e-5, "darcy_K": 1e-3, "forch_CF": 0.55, "inlet_u": -0.03,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": counter_traffic(prec),
                     "algorithmic_bytes_per_launch": algo_bytes, "kernel_ms": round(k_s * 1e3, 4),
                     "pattern_ceiling": pattern_ceiling(k_s, r["stream_only_s"], algo_bytes)},
        "all_site_mlups": round(r["cells"] / k_s / 1e6, 1),
        "reference_published": {"mlups_224": REFERENCE_MLUPS_224, "hardware": "Apple M1/M2 Pro, Metal",
                                "note": "self-inconsistent prose figure, all-site count"},
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
                                       "achieved_GBs": round(76 * r16["active_cells"] / s16 / 1e9, 1),
                                       "frac": round(76 * r16["active_cells"] / s16 / 1e9 / HBM_PEAK_GBS, 4),
                                       "layout": r16["layout"], "traffic": counter_traffic(p16),
                                       "pattern_ceiling": pattern_ceiling(s16, r16["stream_only_s"],
                                                                          76 * r16["active_cells"])}
        log("bench: coupled 224^3 f32 + 50k particles")
        # the particle cloud keeps its 50 000 particles over 300 steps
        out["coupled_224_f32_50k_particles"] = bench_coupled("f32", min(args.steps, 200), min(args.warmup, 100))
        log("bench: adjoint 128^3 f32")
        out["adjoint"] = bench_adjoint()
        log("bench: field diagnostics 224^3 f32")
        out["diagnostics_224_f32"] = bench_diagnostics()
        log("bench: thermal D3Q7 224^3 f32")
        out["thermal_224_f32"] = bench_thermal(steps=args.steps, warmup=args.warmup)
        log("bench: multiphase 224^3 f32")
        out["multiphase_224_f32"] = bench_multiphase(reps=min(args.steps, 100), warmup=args.warmup)
        log("bench: body-force producers 224^3 f32")
        out["forcing_224_f32"] = bench_forcing(reps=min(args.steps, 100), warmup=args.warmup)
        log("bench: main-loop step 224^3 f32")
        out["main_loop_224_f32"] = bench_main_loop(reps=min(args.steps, 50))
        log("bench: cpu baseline")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
