"""The reference's call sequences on the HIP classes, on the device, against the oracle.

- the explicit two-way coupling of LBMSolver.step_with_two_way_coupling
  (legacy/lbm_solver.py:1485-1503: clear_body_force ->
  compute_two_way_coupling_forces(u) -> apply_under_relaxation ->
  add_particle_reaction_forces -> step) with the reference's reaction field
  (coffee_particles.py:1107-1154,1049-1086);
- the protocol verb update_particle_coupling followed by step: one coupling
  acts on one step, a repeated call recomputes, nothing stale remains;
- the plug-in path: create_backend('hip'), UnifiedLBMSolver('hip', ...),
  execute_collision_streaming(adapter, {...}) and (**kwargs), and the field
  views f / u in the reference's [q, x, y, z] / [x, y, z, 3] order;
- enable_les_turbulence switching the eddy viscosity at run time.

Bars: FP32 fluid state within 1e-5 relative L2 of the FP64 oracle; particle
drag within 1e-5, positions within 1e-6.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import rel_l2
from pourover_amd import _lib as L
from pourover_amd import workloads as WL
from pourover_amd.backends import create_backend
from pourover_amd.particles import HipCoffeeParticleSystem, particle_gravity
from pourover_amd.solver import HipLBMSolver, ParticleCoupledLBMProtocol, UnifiedLBMSolver, validate_lbm_solver

pytestmark = pytest.mark.gpu

ALPHA = 0.8


# The explicit reaction is stable only for weak coupling (per-cell drag
# coefficient well below the fluid density): half-size grains and a drag cap of
# 0.05 m keep the 48^3 cup bounded (full-size grains with a cap of 1.0 m
# diverge within ten steps, in the oracle as on the device).
EXPLICIT_RADIUS_SCALE, EXPLICIT_DRAG_CAP = 0.5, 0.05


def _oracle_pp(oracle, wl, count, cap=1.0):
    return oracle.particle_params(count, drag_law=L.POL_DRAG_TEXTBOOK, rho_fluid=1.0,
                                  nu_fluid=(wl.config.tau - 0.5) / 3, drag_max_per_mass=cap, alpha=ALPHA,
                                  g=particle_gravity(1e-5, 1.2), dt=1.0, max_disp=1.0, wall_damping=0.3)


def _initial_velocity(count, seed=3):
    """Particles start moving (float32), so the drag acts from the first step."""
    rng = np.random.default_rng(seed)
    return (rng.normal(scale=0.004, size=(3, count)) + np.array([[0.0], [0.0], [-0.004]])).astype(np.float32)


def _oracle_state(seedp, vel0, radius):
    P = seedp.pos.shape[1]
    return {"pos": seedp.pos.astype(np.float32).astype(np.float64), "vel": vel0.astype(np.float64),
            "radius": radius.astype(np.float32).astype(np.float64),
            "mass": seedp.mass.astype(np.float32).astype(np.float64), "active": np.ones(P, np.int32),
            "drag_old": np.zeros((3, P))}


def test_explicit_reaction_reference_sequence_vs_oracle(hip, oracle):
    """48^3 cup, 1500 particles, 20 steps of step_with_two_way_coupling with the
    reference's explicit reaction field, against the same sequence in the oracle."""
    n, count, steps = 48, 1500, 20
    print("case: explicit reaction, 48^3 cup, %d particles, %d steps" % (count, steps), flush=True)
    wl = WL.v60_config(n, "f32", force_field=True)
    flags = wl.geometry.flags
    seedp = WL.seed_particles(wl.geometry, count)
    solver = HipLBMSolver(wl.config, flags)
    solver.init_fields()
    radius = seedp.radius * EXPLICIT_RADIUS_SCALE
    ps = HipCoffeeParticleSystem(solver.ctx, seedp.pos, radius, seedp.mass, relax_alpha=ALPHA,
                                 gravity=particle_gravity(1e-5, 1.2), reaction="explicit",
                                 drag_max_per_mass=EXPLICIT_DRAG_CAP)
    vel0 = _initial_velocity(count)
    ps.vel.copy_(torch.from_numpy(vel0))
    assert isinstance(solver, ParticleCoupledLBMProtocol)
    for _ in range(steps):
        solver.step_with_two_way_coupling(ps, 1.0, ALPHA)
    torch.cuda.synchronize()

    d = oracle.desc_from_dict(wl.config.oracle_dict(True))
    f = oracle.init_equilibrium(n, n, n)
    pp = _oracle_pp(oracle, wl, count, EXPLICIT_DRAG_CAP)
    st = _oracle_state(seedp, vel0, radius)
    force = np.zeros((3, n, n, n))
    reaction = np.zeros((3, n, n, n))
    for _ in range(steps):
        force[:] = 0.0                                        # clear_body_force
        _, u_field, _ = oracle.moments(d, f, flags, force=force)  # self.u
        reaction[:] = 0.0
        oracle.particles_compute((n, n, n), pp, st, u_field, reaction, explicit=True)
        oracle.particles_relax(st, ALPHA)
        oracle.add_reaction_forces(force, reaction, flags)
        f, _, _, _ = oracle.step(d, f, flags, force=force)
        oracle.particles_advance((n, n, n), pp, st, flags, None)
    rho_o, u_o, _ = oracle.moments(d, f, flags, force=force)

    u_g = solver.u.to_torch().double().cpu().numpy()
    rho_g = solver.rho.to_torch().double().cpu().numpy()
    eu, er = rel_l2(u_g, u_o), rel_l2(rho_g, rho_o)
    react_g = ps.reaction_force_field.to_numpy().transpose(3, 2, 1, 0)
    erx = rel_l2(react_g, reaction)
    act = ps.active.cpu().numpy()
    assert np.array_equal(act, st["active"])
    a = act == 1
    ed = rel_l2(ps.drag_new.double().cpu().numpy()[:, a], st["drag_new"][:, a])
    edr = rel_l2(ps.drag.double().cpu().numpy()[:, a], st["drag"][:, a])
    ep = rel_l2(ps.pos.double().cpu().numpy()[:, a], st["pos"][:, a])
    print("rel L2 u %.2e rho %.2e reaction %.2e drag_new %.2e drag %.2e pos %.2e" % (eu, er, erx, ed, edr, ep))
    assert np.abs(u_o).max() < 0.05  # the weak coupling stayed bounded
    assert np.abs(st["drag_new"]).max() > 1e-4  # and acted
    assert eu < 1e-5 and er < 1e-6
    assert erx < 1e-5 and ed < 1e-5 and edr < 1e-5
    assert ep < 1e-6
    # the body force holds exactly the last reaction on the fluid cells (gravity is the constant G)
    bf = solver.body_force.to_torch().double().cpu().numpy()
    fo = force.copy()
    assert rel_l2(bf, fo) < 1e-5
    # Newton's third law on the device field: sum of the reaction = -sum of drag_new
    tot = react_g.reshape(3, -1).sum(axis=1)
    dsum = ps.drag_new.double().cpu().numpy()[:, a].sum(axis=1)
    np.testing.assert_allclose(tot, -dsum, rtol=1e-4, atol=1e-12)
    diag = ps.get_coupling_diagnostics()
    assert diag["active_particles"] == int(a.sum())
    assert diag["max_reaction_force"] == pytest.approx(float(np.linalg.norm(react_g, axis=0).max()), rel=1e-6)
    stats = ps.get_particle_statistics()
    assert stats["count"] == int(a.sum()) and stats["invalid_count"] == 0
    assert stats["positions"].shape == (int(a.sum()), 3)


def test_update_particle_coupling_then_step_is_the_coupled_step(hip):
    """Implicit reaction, first step from the primed state: update_particle_coupling
    + step + advance leaves the particles bit for bit, and the populations to
    FP32 round-off, where the native coupled step does; a repeated
    update_particle_coupling recomputes instead of adding twice; after the
    step no drag is left in the field."""
    wl = WL.v60_config(40, "f32", particle_field=True)
    flags = wl.geometry.flags
    seedp = WL.seed_particles(wl.geometry, 2000)
    print("case: update_particle_coupling vs coupled_step, 40^3 cup, 2000 particles", flush=True)
    outs = []
    for mode in ("verbs", "verbs_twice", "native"):
        solver = HipLBMSolver(wl.config, flags)
        solver.init_fields()
        ps = HipCoffeeParticleSystem(solver.ctx, seedp.pos, seedp.radius, seedp.mass, relax_alpha=ALPHA,
                                     gravity=particle_gravity(1e-5, 1.2))
        ps.vel.copy_(torch.from_numpy(_initial_velocity(2000)))
        ps.prime()
        if mode == "native":
            solver.step_with_two_way_coupling(ps, 1.0, ALPHA)
        else:
            solver.update_particle_coupling(ps)
            if mode == "verbs_twice":
                solver.update_particle_coupling(ps)
            assert float(solver.ctx.pdrag[3].max()) > 0.0  # the coupling spread drag
            solver.step()
            assert float(solver.ctx.pdrag.abs().max()) == 0.0  # consumed by exactly this step
            ps.advance()
        torch.cuda.synchronize()
        outs.append((solver.ctx.f_current.clone(), ps.pos.clone(), ps.vel.clone(), ps.drag_new.clone()))
    a, b, c = outs
    # drag_new, and with it the particle push, does not depend on the order of
    # the atomic additions into the drag field: bit for bit
    for x, y in zip(a[1:], c[1:]):
        assert torch.equal(x.view(torch.int32), y.view(torch.int32))
    assert torch.equal(a[3].view(torch.int32), b[3].view(torch.int32))
    # the populations read the drag field, whose FP32 atomic sums may round in
    # another order: equal to FP32 round-off (a second, stale coupling would
    # double the particle force and differ at ~1e-3)
    assert rel_l2(a[0].double().cpu().numpy(), c[0].double().cpu().numpy()) < 1e-6
    assert rel_l2(a[0].double().cpu().numpy(), b[0].double().cpu().numpy()) < 1e-6


def test_plugin_path_unified_solver_vs_oracle(hip, oracle):
    """create_backend('hip') + UnifiedLBMSolver('hip', cfg, flags) on the 40^3 cup:
    validate_lbm_solver, execute_collision_streaming(adapter, {'tau': 0.6}) and
    the (**kwargs) form, then u (NX, NY, NZ, 3) and raw f (19, NX, NY, NZ)
    through the field views against the oracle."""
    wl = WL.v60_config(40, "f32")
    flags = wl.geometry.flags
    print("case: plug-in path, 40^3 cup", flush=True)
    be = create_backend("hip")
    assert be.validate_platform()
    solver = UnifiedLBMSolver("hip", wl.config, flags)
    assert validate_lbm_solver(solver)
    solver.init_fields()
    adapter = solver.memory_adapter
    steps = 0
    for _ in range(10):
        solver.backend.execute_collision_streaming(adapter, {"tau": 0.6})
        steps += 1
    for _ in range(10):
        solver.backend.execute_collision_streaming(adapter, tau=0.6)
        steps += 1
    solver.backend.execute_collision_streaming(adapter, {"tau": 0.6, "steps": 5})
    steps += 5
    metrics = solver.backend.get_performance_metrics()
    for k in ("collision_time", "streaming_time", "boundary_time", "total_time", "throughput_mlups"):
        assert k in metrics
    u_ref_order = solver.u.to_numpy()
    f_ref_order = solver.f.to_numpy()
    assert u_ref_order.shape == (40, 40, 40, 3)
    assert f_ref_order.shape == (19, 40, 40, 40)

    d = oracle.desc_from_dict(wl.config.oracle_dict(True))
    f = oracle.init_equilibrium(40, 40, 40)
    for _ in range(steps):
        f, _, _, _ = oracle.step(d, f, flags)
    rho_o, u_o, _ = oracle.moments(d, f, flags)
    u_o_ref = u_o.transpose(3, 2, 1, 0)  # [3,z,y,x] -> [x,y,z,3]
    assert rel_l2(u_ref_order, u_o_ref) < 1e-5
    fluid = (((flags >> L.POL_KIND_SHIFT) & L.POL_KIND_MASK) != L.POL_KIND_SOLID)
    fo = f.transpose(0, 3, 2, 1)  # [19,z,y,x] -> [19,x,y,z]
    m = fluid.transpose(2, 1, 0)
    assert rel_l2(f_ref_order[:, m], fo[:, m]) < 1e-6
    assert rel_l2(solver.rho.to_numpy(), rho_o.transpose(2, 1, 0)) < 1e-6
    # the other views: u components and the solid mask in reference order
    assert np.array_equal(solver.ux.to_numpy(), u_ref_order[..., 0])
    assert np.array_equal(solver.solid.to_numpy(), (~m).astype(np.uint8))


def test_enable_les_turbulence_switches_the_eddy_viscosity(hip, oracle):
    """A solver built without LES, switched on by enable_les_turbulence(0.17)
    after 30 steps and off again after 60, against the oracle doing the same."""
    n = 40
    wl_off = WL.v60_config(n, "f32", les=False)
    wl_on = WL.v60_config(n, "f32", les=True)
    flags = wl_on.geometry.flags
    print("case: LES switched at run time, 40^3 cup", flush=True)
    solver = HipLBMSolver(wl_off.config, flags)
    solver.init_fields()
    solver.run(30)
    solver.enable_les_turbulence(0.17)
    assert solver.use_les and solver.config.les
    solver.run(30)
    solver.disable_les_turbulence()
    solver.run(20)
    u_g = solver.u.to_torch().double().cpu().numpy()

    f = oracle.init_equilibrium(n, n, n)
    for cfg, k in ((wl_off.config, 30), (wl_on.config, 30), (wl_off.config, 20)):
        d = oracle.desc_from_dict(cfg.oracle_dict(True))
        for _ in range(k):
            f, _, _, _ = oracle.step(d, f, flags)
    _, u_o, _ = oracle.moments(d, f, flags)
    assert rel_l2(u_g, u_o) < 1e-5
    # and LES did act: the run without the switch differs
    g = oracle.init_equilibrium(n, n, n)
    d_off = oracle.desc_from_dict(wl_off.config.oracle_dict(True))
    for _ in range(80):
        g, _, _, _ = oracle.step(d_off, g, flags)
    _, u_plain, _ = oracle.moments(d_off, g, flags)
    assert rel_l2(u_plain, u_o) > 1e-4
