"""Precision forms of the V60 run: populations stored as scaled deviations
D = 36 (f - w) in FP16, BF16 and FP32, collision and particle arithmetic in
FP32, reductions in FP64, against an FP64 run of the same seeded inputs
(224^3 cup, 2000 steps; with and without the 50 000 coupled particles).

Reported per storage: relative L2 error of velocity, density and (coupled
run) particle position after 2000 steps.  Bars: FP32 1e-5 on u for the fluid
alone; FP16 1e-3 on u (the target for 16-bit storage); BF16 (8-bit
mantissa) with a loose sanity bound.  The coupled run moves particles across
cell faces and wall-rejection thresholds, which turns round-off differences
into different particle paths, so its FP32 bar is 1e-4.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import rel_l2
from pourover_amd import workloads as WL
from pourover_amd.engine import HipContext
from pourover_amd.particles import HipCoffeeParticleSystem, particle_gravity

pytestmark = pytest.mark.gpu

N, COUNT, STEPS = 224, 50_000, 2000


def _run(precision: str, particles: bool):
    wl = WL.v60_config(N, precision, particle_field=particles)
    ctx = HipContext(wl.config, wl.geometry.flags)
    ctx.init_equilibrium()
    out = {"solid": wl.geometry.kind == 1}
    if particles:
        seedp = WL.seed_particles(wl.geometry, COUNT)
        ps = HipCoffeeParticleSystem(ctx, seedp.pos, seedp.radius, seedp.mass, gravity=particle_gravity(1e-5, 1.2))
        ps.prime()
        ps.coupled_step(nsteps=STEPS)
        out["pos"] = ps.pos.double().cpu().numpy()
        out["active"] = ps.active.cpu().numpy().astype(bool)
        del ps
    else:
        ctx.step(STEPS)
    ctx.macroscopic(macro=True)
    torch.cuda.synchronize()
    out["u"] = ctx.u.double().cpu().numpy()
    out["rho"] = ctx.rho.double().cpu().numpy()
    del ctx
    torch.cuda.empty_cache()
    return out


_REF = {}


def _reference(particles: bool):
    if particles not in _REF:
        print("case: precision reference f64 224^3, particles=%s, %d steps" % (particles, STEPS), flush=True)
        _REF[particles] = _run("f64", particles)
    return _REF[particles]


def _errors(r, ref):
    fluid = ~ref["solid"]
    eu = rel_l2(r["u"][:, fluid], ref["u"][:, fluid])
    er = rel_l2(r["rho"][fluid], ref["rho"][fluid])
    erd = rel_l2(r["rho"][fluid] - 1.0, ref["rho"][fluid] - 1.0)
    return eu, er, erd


@pytest.mark.parametrize("precision,u_bar", [("f32", 1e-5), ("f16", 1e-3), ("bf16", 5e-2)])
def test_fluid_storage_precision(precision, u_bar):
    ref = _reference(False)
    print("case: precision %s fluid only" % precision, flush=True)
    r = _run(precision, False)
    eu, er, erd = _errors(r, ref)
    print("fluid %s: rel L2 u %.3e  rho %.3e  (rho - 1 %.3e)" % (precision, eu, er, erd))
    assert np.isfinite(r["u"]).all()
    assert eu < u_bar
    assert er < 1e-5


@pytest.mark.parametrize("precision,u_bar", [("f32", 1e-4), ("f16", 1e-3), ("bf16", 5e-2)])
def test_coupled_storage_precision(precision, u_bar):
    ref = _reference(True)
    print("case: precision %s coupled" % precision, flush=True)
    r = _run(precision, True)
    eu, er, erd = _errors(r, ref)
    both = r["active"] & ref["active"]
    ep = rel_l2(r["pos"][:, both], ref["pos"][:, both])
    print("coupled %s: rel L2 u %.3e  rho %.3e  (rho - 1 %.3e)  particle pos %.3e  (active %d / %d)"
          % (precision, eu, er, erd, ep, int(both.sum()), int(ref["active"].sum())))
    assert np.isfinite(r["u"]).all()
    assert eu < u_bar
    assert er < 1e-5
    assert ep < 1e-2
