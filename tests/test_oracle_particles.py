"""Pins of the oracle's explicit two-way coupling (the reference's reaction
field, coffee_particles.py:1107-1154,1049-1086), its under-relaxation
(:1200-1212) and the body-force injection (legacy/lbm_solver.py:1478-1483).

Each check is a closed form or an invariant computed here independently of
the oracle's code: Newton's third law over the whole field, the 8-corner
trilinear spread of one particle in a uniform flow with the drag law written
out by hand, the relaxation formula, and the solid mask of the injection.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from pourover_amd import _lib as L


def _params(oracle, P, law, alpha=0.8, **kw):
    base = dict(drag_law=law, rho_fluid=1.0, nu_fluid=(0.6 - 0.5) / 3.0, drag_max_per_mass=100.0, alpha=alpha,
                g=(0.0, 0.0, 0.0), dt=1.0, max_disp=1.0, wall_damping=0.3)
    base.update(kw)
    return oracle.particle_params(P, **base)


def _state(pos, radius, mass, vel=None):
    P = pos.shape[1]
    return {"pos": np.ascontiguousarray(pos, dtype=np.float64),
            "vel": np.zeros((3, P)) if vel is None else np.ascontiguousarray(vel, dtype=np.float64),
            "radius": np.ascontiguousarray(radius, dtype=np.float64),
            "mass": np.ascontiguousarray(mass, dtype=np.float64), "active": np.ones(P, np.int32),
            "drag_old": np.zeros((3, P))}


def _schiller_naumann(re):
    if re < 0.1:
        return 24.0 / max(0.01, re)
    if re < 1000.0:
        return 24.0 / re * (1.0 + 0.15 * re ** 0.687)
    return 0.44


@pytest.mark.parametrize("law", [L.POL_DRAG_TEXTBOOK, L.POL_DRAG_REFERENCE])
def test_single_particle_reaction_closed_form(oracle, law):
    """One particle at rest in a uniform flow: drag_new = D u0/|u0| with
    D = 1/2 rho C_D pi' r^2 |u0| (x |u0| for the textbook law), pi' = 3.14159
    as the reference writes it, and corner (i+a, j+b, k+c) of the reaction
    field holds -w_abc drag_new with the trilinear weights of the fractions;
    every other cell stays zero."""
    n = 10
    dims = (n, n, n)
    u0 = np.array([0.011, -0.023, 0.0045])
    u = np.empty((3, n, n, n))
    u[:] = u0[:, None, None, None]
    pos = np.array([[3.3], [4.7], [5.25]])
    r = 0.9
    pp = _params(oracle, 1, law)
    st = _state(pos, np.array([r]), np.array([50.0]))
    field = np.zeros((3, n, n, n))
    oracle.particles_compute(dims, pp, st, u, field, explicit=True)
    mag = float(np.linalg.norm(u0))
    mu = 1.0 * (0.6 - 0.5) / 3.0
    re = 1.0 * mag * 2 * r / mu
    cd = _schiller_naumann(re)
    dmag = 0.5 * 1.0 * cd * 3.14159 * r * r * mag
    if law == L.POL_DRAG_TEXTBOOK:
        dmag *= mag
    drag = dmag * u0 / mag
    np.testing.assert_allclose(st["drag_new"][:, 0], drag, rtol=1e-13)
    assert st["re_p"][0] == pytest.approx(re, rel=1e-14)
    assert st["cd"][0] == pytest.approx(cd, rel=1e-13)
    np.testing.assert_allclose(st["u_fluid"][:, 0], u0, rtol=1e-15)
    fx, fy, fz = 0.3, 0.7, 0.25
    expect = np.zeros_like(field)
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                w = (fx if a else 1 - fx) * (fy if b else 1 - fy) * (fz if c else 1 - fz)
                expect[:, 5 + c, 4 + b, 3 + a] = -w * drag
    np.testing.assert_allclose(field, expect, rtol=1e-12, atol=1e-18)
    # relaxation is not part of compute_two_way_coupling_forces
    assert "drag" not in st


def test_newton_third_law_over_the_field(oracle):
    """Sum over cells of the reaction = -sum over particles of drag_new (FP64):
    the 8 weights of every particle sum to one, whatever its position."""
    n = 24
    dims = (n, n, n)
    rng = np.random.default_rng(5)
    P = 3000
    pos = rng.uniform(-1.0, n + 1.0, size=(3, P))  # includes clamped positions past the faces
    radius = rng.uniform(0.2, 1.5, size=P)
    mass = 1.2 * 4.0 / 3.0 * math.pi * radius ** 3
    vel = rng.normal(scale=0.01, size=(3, P))
    u = rng.normal(scale=0.02, size=(3, n, n, n))
    pp = _params(oracle, P, L.POL_DRAG_TEXTBOOK, drag_max_per_mass=1.0)
    st = _state(pos, radius, mass, vel)
    st["active"][::7] = 0
    field = np.zeros((3, n, n, n))
    oracle.particles_compute(dims, pp, st, u, field, explicit=True)
    total = field.reshape(3, -1).sum(axis=1)
    act = st["active"].astype(bool)
    drag_sum = st["drag_new"][:, act].sum(axis=1)
    assert np.all(np.abs(drag_sum) > 0)
    np.testing.assert_allclose(total, -drag_sum, rtol=1e-12)
    # inactive particles produce nothing and keep zero outputs
    assert np.all(st["drag_new"][:, ~act] == 0)


def test_under_relaxation_closed_form(oracle):
    """drag = a drag_new + (1 - a) drag_old; drag_old <- drag; inactive untouched."""
    rng = np.random.default_rng(2)
    P = 50
    st = {"active": np.ones(P, np.int32), "drag_new": rng.normal(size=(3, P)), "drag_old": rng.normal(size=(3, P))}
    st["active"][3] = 0
    old = st["drag_old"].copy()
    a = 0.8
    oracle.particles_relax(st, a)
    act = st["active"].astype(bool)
    want = a * st["drag_new"] + (1 - a) * old
    np.testing.assert_allclose(st["drag"][:, act], want[:, act], rtol=1e-15)
    np.testing.assert_array_equal(st["drag_old"][:, act], st["drag"][:, act])
    np.testing.assert_array_equal(st["drag_old"][:, ~act], old[:, ~act])
    assert np.all(st["drag"][:, ~act] == 0)
    # a = 1 keeps the new value, a = 0 the history
    st2 = {"active": np.ones(P, np.int32), "drag_new": st["drag_new"].copy(), "drag_old": old.copy()}
    oracle.particles_relax(st2, 1.0)
    np.testing.assert_array_equal(st2["drag"], st["drag_new"])


def test_reaction_injection_skips_solid_cells(oracle):
    """legacy/lbm_solver.py:1478-1483: body_force += reaction where solid == 0."""
    n = 6
    rng = np.random.default_rng(1)
    kind = rng.integers(0, 4, size=(n, n, n)).astype(np.uint32)
    flags = kind << L.POL_KIND_SHIFT
    reaction = rng.normal(size=(3, n, n, n))
    force = rng.normal(size=(3, n, n, n))
    before = force.copy()
    oracle.add_reaction_forces(force, reaction, flags)
    solid = kind == L.POL_KIND_SOLID
    np.testing.assert_array_equal(force[:, solid], before[:, solid])
    np.testing.assert_array_equal(force[:, ~solid], before[:, ~solid] + reaction[:, ~solid])


def test_implicit_compute_matches_relaxed_couple_field(oracle):
    """The implicit form of compute_two_way_coupling_forces writes the same drag
    field and drag_new as the relaxed coupling pass (relaxation only touches
    drag / drag_old)."""
    n = 16
    dims = (n, n, n)
    rng = np.random.default_rng(9)
    P = 400
    pos = rng.uniform(0.5, n - 1.5, size=(3, P))
    radius = rng.uniform(0.3, 1.0, size=P)
    mass = 1.2 * 4.0 / 3.0 * math.pi * radius ** 3
    u = rng.normal(scale=0.02, size=(3, n, n, n))
    pp = _params(oracle, P, L.POL_DRAG_TEXTBOOK, drag_max_per_mass=1.0)
    a, b = _state(pos, radius, mass), _state(pos, radius, mass)
    fa, fb = np.zeros((4, n, n, n)), np.zeros((4, n, n, n))
    oracle.particles_compute(dims, pp, a, u, fa, explicit=False)
    oracle.particles_couple(dims, pp, b, u, fb)
    np.testing.assert_array_equal(fa, fb)
    np.testing.assert_array_equal(a["drag_new"], b["drag_new"])
    # the drag field's B component is the sum of the weights times beta: positive where touched
    assert np.all(fa[3] >= 0) and fa[3].sum() > 0
