"""GPU parity of the particle coupling kernels (pol_particle_index,
pol_particles_couple/advance) against the C oracle, through the C ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from conftest import rel_l2
from pourover_amd import _lib as L
from pourover_amd import workloads as WL
from pourover_amd.engine import HipContext
from pourover_amd.particles import HipCoffeeParticleSystem, particle_gravity

pytestmark = pytest.mark.gpu


def _edge_positions(n, count=200_000, seed=11):
    rng = np.random.default_rng(seed)
    f32 = np.float32
    special = []
    for k in (0, 1, 2, n // 2, n - 3, n - 2, n - 1, n):
        k = f32(k)
        special += [k, np.nextafter(k, f32(-1e9)), np.nextafter(k, f32(1e9)), k + f32(0.5)]
    special += [f32(-0.0), f32(-1e-30), f32(-5), f32(3 * n), f32(np.inf), f32(-np.inf), f32(np.nan)]
    special = np.array(special, dtype=f32)
    rnd = rng.uniform(-2, n + 2, size=count).astype(f32)
    vals = np.concatenate([special, rnd])
    return np.stack([vals, np.roll(vals, 7), np.roll(vals, 13)]).astype(np.float32)


@pytest.mark.parametrize("dims", [(16, 16, 16), (224, 224, 224), (37, 5, 64)])
def test_particle_index_bit_exact(hip, oracle, dims):
    """coffee_particles.py:1159-1174 index and clamped fractions: bit-exact."""
    pos = _edge_positions(max(dims))
    n = pos.shape[1]
    dpos = torch.from_numpy(pos).cuda()
    idx = torch.empty((3, n), dtype=torch.int32, device="cuda")
    frac = torch.empty((3, n), dtype=torch.float32, device="cuda")
    st = hip.pol_particle_index(L.ptr(dpos), n, dims[0], dims[1], dims[2], L.ptr(idx), L.ptr(frac), L.stream_handle())
    assert st == L.POL_OK
    torch.cuda.synchronize()
    oi, of = oracle.particle_index(pos, dims)
    assert np.array_equal(idx.cpu().numpy(), oi)
    assert np.array_equal(frac.cpu().numpy().view(np.uint32), of.view(np.uint32))


def _setup(n, precision, count, seed=WL.SEED):
    wl = WL.v60_config(n, precision, particle_field=True)
    seedp = WL.seed_particles(wl.geometry, count, seed=seed)
    return wl, seedp


def _oracle_coupled(oracle, wl, seedp, steps, alpha=0.8):
    cfg = wl.config
    n = cfg.nx
    flags = wl.geometry.flags
    d = oracle.desc_from_dict(cfg.oracle_dict(True))
    f = oracle.init_equilibrium(n, n, n)
    _, u_field, _ = oracle.moments(d, f, flags)
    P = seedp.pos.shape[1]
    g = particle_gravity(1e-5, 1.2)
    pp = oracle.particle_params(P, drag_law=L.POL_DRAG_TEXTBOOK, rho_fluid=1.0, nu_fluid=(cfg.tau - 0.5) / 3,
                                drag_max_per_mass=1.0, alpha=alpha, g=g, dt=1.0, max_disp=1.0, wall_damping=0.3)
    st = {"pos": seedp.pos.astype(np.float32).astype(np.float64), "vel": np.zeros((3, P)),
          "radius": seedp.radius.astype(np.float32).astype(np.float64),
          "mass": seedp.mass.astype(np.float32).astype(np.float64), "active": np.ones(P, np.int32),
          "drag_old": np.zeros((3, P))}
    field = np.zeros((4, n, n, n))
    for _ in range(steps):
        oracle.particles_couple((n, n, n), pp, st, u_field, field)
        f, rho, u_field, _ = oracle.step(d, f, flags, want_macro=True, pdrag=field)
        oracle.particles_advance((n, n, n), pp, st, flags, field)
    return rho, u_field, st


def _gpu_coupled(wl, seedp, steps, alpha=0.8):
    ctx = HipContext(wl.config, wl.geometry.flags)
    ctx.init_equilibrium()
    ps = HipCoffeeParticleSystem(ctx, seedp.pos, seedp.radius, seedp.mass, relax_alpha=alpha,
                                 gravity=particle_gravity(1e-5, 1.2))
    ps.prime()
    ps.coupled_step(nsteps=steps)
    torch.cuda.synchronize()
    return ctx, ps


def test_coupled_steps_match_oracle(hip, oracle):
    """48^3 cup, 3000 particles, 60 coupled steps: fluid and particle state vs the FP64 oracle."""
    wl, seedp = _setup(48, "f32", 3000)
    print("case: particles 48^3 x 3000, 60 steps", flush=True)
    ctx, ps = _gpu_coupled(wl, seedp, 60)
    rho_o, u_o, st = _oracle_coupled(oracle, wl, seedp, 60)
    eu = rel_l2(ctx.u.double().cpu().numpy(), u_o)
    er = rel_l2(ctx.rho.double().cpu().numpy(), rho_o)
    act = ps.active.cpu().numpy()
    assert np.array_equal(act, st["active"])
    ep = rel_l2(ps.pos.double().cpu().numpy()[:, act == 1], st["pos"][:, act == 1])
    ev = rel_l2(ps.vel.double().cpu().numpy()[:, act == 1], st["vel"][:, act == 1])
    print("rel L2 u %.2e rho %.2e pos %.2e vel %.2e" % (eu, er, ep, ev))
    assert eu < 1e-5 and er < 1e-6
    assert ep < 1e-6 and ev < 1e-4


def test_momentum_exchange_and_footprint_clear(hip):
    """The drag-field scatter is a partition of unity (FP64 field sums equal the
    particle sums of beta v and beta), and the field is all zero after advance
    (the 8-corner footprint is cleared)."""
    wl, seedp = _setup(64, "f32", 5000)
    ctx = HipContext(wl.config, wl.geometry.flags)
    ctx.init_equilibrium()
    ps = HipCoffeeParticleSystem(ctx, seedp.pos, seedp.radius, seedp.mass, gravity=particle_gravity(1e-5, 1.2))
    ps.prime()
    ps.coupled_step(nsteps=30)
    ps.couple()
    field, parts = ps.drag_field_sums()
    assert np.allclose(field, parts, rtol=1e-5, atol=1e-12)
    assert parts[3] > 0
    ctx.step(1, L.POL_WANT_MACRO)
    ps.advance()
    torch.cuda.synchronize()
    assert float(ctx.pdrag.abs().max()) == 0.0


def test_full_size_coupled_run_stays_physical(hip):
    """224^3 cup with the 50 000-particle seed, 200 coupled steps: finite state,
    particles inside the cup, none inside walls or the filter."""
    wl, seedp = _setup(224, "f32", 50_000)
    ctx, ps = _gpu_coupled(wl, seedp, 200)
    red = ctx.reduce()
    assert red["nonfinite_cells"] == 0 and red["max_velocity"] < 0.3
    pos = ps.pos.cpu().numpy()
    act = ps.active.cpu().numpy() == 1
    assert act.sum() > 0.9 * 50_000
    ci = pos[:, act].astype(np.int64)
    kind = wl.geometry.kind[ci[2], ci[1], ci[0]]
    assert (kind == 0).all()
    assert not wl.geometry.porous[ci[2], ci[1], ci[0]].any()
    assert np.isfinite(pos).all()


def test_coupled_step_without_particles_is_the_plain_step(hip):
    """An empty particle set (count 0): the coupled loop launches no particle
    work, the drag field stays zero, and the fluid ends bit for bit where the
    plain step takes it."""
    print("case: coupled step with 0 particles, 32^3 cup, 20 steps", flush=True)
    wl = WL.v60_config(32, "f32", particle_field=True)
    a = HipContext(wl.config, wl.geometry.flags)
    b = HipContext(wl.config, wl.geometry.flags)
    a.init_equilibrium()
    b.init_equilibrium()
    none = np.zeros(0, np.float32)
    ps = HipCoffeeParticleSystem(a, np.zeros((3, 0), np.float32), none, none, gravity=particle_gravity(1e-5, 1.2))
    ps.prime()
    ps.coupled_step(nsteps=20)
    b.step(20, L.POL_WANT_MACRO)
    torch.cuda.synchronize()
    assert a.current == b.current
    assert torch.equal(a.f_current.view(torch.int32), b.f_current.view(torch.int32))
    assert torch.equal(a.rho.view(torch.int32), b.rho.view(torch.int32))
    assert torch.equal(a.u.view(torch.int32), b.u.view(torch.int32))
    assert float(a.pdrag.abs().max()) == 0.0
    assert ps.get_coupling_diagnostics()["active_particles"] == 0
