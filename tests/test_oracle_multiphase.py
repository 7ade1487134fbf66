"""Pins of the multiphase oracle (oracle/multiphase.py) by closed forms:
planar and spherical interfaces (normal, curvature = -2/R), exact
Laplacians of quadratics, the upwind direction, the phase-separation
increment, the density/phase map and the surface-tension guards."""
from __future__ import annotations

import numpy as np

from oracle import multiphase as OM

F = np.float32
N = 20


def _state(phi, u=None, rho=None, solid=None):
    shape = phi.shape
    return OM.new_state(phi, np.zeros(shape + (3,), F) if u is None else u,
                        np.ones(shape, F) if rho is None else rho,
                        np.zeros(shape, np.uint8) if solid is None else solid)


def _grid(n=N):
    return np.meshgrid(np.arange(n, dtype=F), np.arange(n, dtype=F), np.arange(n, dtype=F), indexing="ij")


def test_reference_constants():
    assert abs(OM.SURFACE_TENSION_LU - 2.4 ** 2 * 0.000625) < 1e-15
    assert abs(OM.RHO_AIR - 1.204 / 965.3) < 1e-18


def test_planar_interface_normal_and_zero_curvature():
    x, y, z = _grid()
    phi = np.tanh((z - F(9.5)) / F(2)).astype(F)
    s = _state(phi)
    OM.gradients(s)
    OM.curvature(s)
    OM.surface_tension_force(s)
    n = s["normal"][1:-1, 1:-1, 1:-1]
    assert np.all(n[..., 0] == 0) and np.all(n[..., 1] == 0) and np.all(n[..., 2] == 1)
    k = s["curvature"]
    assert np.all(k[1:-1, 1:-1, 2:-2] == 0)          # interior neighbours only
    assert np.all(k[1:-1, 1:-1, 1] != 0)             # the face neighbour's normal is 0
    assert np.all(s["surface_force"][1:-1, 1:-1, 2:-2] == 0)


def test_sphere_curvature():
    n = 40
    x, y, z = _grid(n)
    R = 12.0
    r = np.sqrt((x - 19.5) ** 2 + (y - 19.5) ** 2 + (z - 19.5) ** 2)
    phi = np.tanh((R - r) / 2.0).astype(F)           # liquid (+1) inside
    s = _state(phi)
    OM.gradients(s)
    OM.curvature(s)
    OM.surface_tension_force(s)
    band = (np.abs(r - R) < 0.7)
    band[0, :, :] = band[-1, :, :] = band[:, 0, :] = band[:, -1, :] = band[:, :, 0] = band[:, :, -1] = False
    k = s["curvature"][band].astype(np.float64)
    # n = grad phi / |grad phi| points inward: div n = -2 / r
    np.testing.assert_allclose(np.mean(k * r[band]), -2.0, rtol=0.03)
    # CSF force points along n with magnitude sigma |kappa| |grad phi|, i.e. inward-normal times a negative
    f = s["surface_force"][band].astype(np.float64)
    nn = s["normal"][band].astype(np.float64)
    assert np.all(np.sum(f * nn, axis=1) < 0)


def test_laplacian_and_chemical_potential_closed_form():
    x, y, z = _grid()
    phi = (F(0.25) * x * x + F(0.5) * y * y - F(0.125) * z * z).astype(F) / F(1024)
    s = _state(phi)
    OM.chemical_potential(s)
    lap = s["laplacian_phi"][1:-1, 1:-1, 1:-1]
    assert np.all(lap == F(2 * (0.25 + 0.5 - 0.125) / 1024))
    c = _state(np.full((8, 8, 8), F(0.5)))
    OM.chemical_potential(c)
    assert np.all(c["mu"][1:-1, 1:-1, 1:-1] == F(0.125) - F(0.5))
    assert np.all(c["mu"][0] == 0)                     # faces untouched


def test_upwind_direction():
    x, y, z = _grid()
    phi = (F(1e-3) * x * x).astype(F)
    for ux in (0.05, -0.05):
        u = np.zeros((N, N, N, 3), F)
        u[..., 0] = ux
        s = _state(phi, u=u)
        OM.update_phase_field(s)
        xi = x[1:-1, 1:-1, 1:-1].astype(np.float64)
        d = 1e-3 * (2 * xi - 1) if ux > 0 else 1e-3 * (2 * xi + 1)
        want = np.clip(phi[1:-1, 1:-1, 1:-1] - ux * d, -1, 1)
        np.testing.assert_allclose(s["phi_new"][1:-1, 1:-1, 1:-1], want, rtol=2e-6, atol=1e-9)


def test_mobility_term_on_quadratic_mu():
    x, y, z = _grid()
    mu = (F(0.5) * z * z).astype(F)
    s = _state(np.zeros((N, N, N), F))
    s["mu"] = mu
    OM.update_phase_field(s)
    assert np.all(s["phi_new"][1:-1, 1:-1, 1:-1] == F(0.001) * F(1.0))  # M * lap(mu) = 0.001 * 1


def test_phase_separation_increment():
    for p, changed in ((0.5, True), (-0.3, True), (0.995, False), (-1.0, False)):
        s = _state(np.full((6, 6, 6), F(p)))
        OM.update_phase_field(s)
        OM.phase_separation(s)
        got = s["phi_new"][2, 2, 2]
        want = p - 0.001 * p * (p * p - 1.0) if changed else p
        assert abs(float(got) - want) < 1e-7, (p, got, want)


def test_density_and_phase_map():
    phi = np.array([-1.0, 1.0, 0.0, 2.0, -3.0], F).reshape(5, 1, 1)
    s = {"phi": phi}
    OM.density_from_phase(s)
    rho, ph = s["rho"].ravel(), s["phase"].ravel()
    assert rho[0] == F(OM.RHO_AIR) and ph[0] == 0
    assert abs(float(rho[1]) - 1.0) < 1e-7 and ph[1] == 1
    assert abs(float(rho[2]) - (1.0 + OM.RHO_AIR) / 2) < 1e-7 and ph[2] == F(0.5)
    assert rho[3] == rho[1] and rho[4] == rho[0]


def test_surface_tension_guards():
    x, y, z = _grid()
    r = np.sqrt((x - 9.5) ** 2 + (y - 9.5) ** 2 + (z - 9.5) ** 2)
    phi = np.tanh((6.0 - r) / 2.0).astype(F)
    solid = np.zeros(phi.shape, np.uint8)
    solid[:10] = 1
    rho = np.ones(phi.shape, F)
    rho[:, :10] = F(1e-12)
    s = _state(phi, rho=rho, solid=solid)
    OM.accumulate_surface_tension_pre_collision(s)
    bf = s["body_force"]
    f = s["surface_force"]
    assert np.all(bf[:10] == 0) and np.all(bf[:, :10] == 0)
    sel = np.zeros(phi.shape, bool)
    sel[10:, 10:] = True
    np.testing.assert_array_equal(bf[sel], f[sel] / rho[sel][:, None])
    assert np.abs(f).max() > 0
    assert np.all(f[np.abs(phi) >= 0.9] == 0)


def test_step_sequence_matches_parts():
    rng = np.random.default_rng(0)
    x, y, z = _grid()
    r = np.sqrt((x - 9.5) ** 2 + (y - 9.5) ** 2 + (z - 9.5) ** 2)
    phi = np.tanh((6.0 - r) / 2.0).astype(F)
    u = (rng.random((N, N, N, 3), dtype=F) - F(0.5)) * F(0.02)
    s = _state(phi, u=u)
    t = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in s.items()}
    OM.step(s, step_count=20)
    OM.gradients(t)
    OM.curvature(t)
    OM.surface_tension_force(t)
    OM.apply_surface_tension(t)
    OM.update_phase_field(t)
    OM.phase_separation(t)
    t["phi"] = t["phi_new"].copy()
    OM.density_from_phase(t)
    for k in ("phi", "rho", "phase", "body_force"):
        np.testing.assert_array_equal(s[k], t[k])
    assert np.all(s["phi"][0] == 0)                   # faces take phi_new's (never written) values
    st = OM.interface_statistics(s)
    assert 0 < st["water_fraction"] < 1 and st["max_curvature"] > 0
