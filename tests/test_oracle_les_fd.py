"""Pins of the oracle's finite-difference eddy viscosity (oracle/les_fd.py,
les_turbulence.py:318-380) and of its use in the oracle step (tau_eff =
tau + 3 nu, clamped to [0.55, 1.90], legacy/lbm_solver.py:565-572).

Each check is a closed form fixed independently of the restatement: a
uniform shear gives |S| = gamma exactly, a rigid rotation has no strain, the
thresholds, the clamp, the faces and the masks of the reference; and a
given viscosity collides exactly like a molecular tau of tau + 3 nu.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle.les_fd import les_fd_viscosity, mask_from_flags
from pourover_amd import _lib as L

N = 10


def _grid():
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    return x.astype(np.float64), y.astype(np.float64), z.astype(np.float64)


@pytest.mark.parametrize("gamma", [2e-3, 0.01, 0.05])
def test_uniform_shear_closed_form(gamma):
    """u = (gamma z, 0, 0): S13 = gamma / 2, |S| = sqrt(2 * 2 (gamma/2)^2) = gamma,
    nu = C_s^2 gamma on every interior cell."""
    x, y, z = _grid()
    u = np.zeros((3, N, N, N))
    u[0] = gamma * z
    cs = 0.17
    nu = les_fd_viscosity(u, cs=cs, dtype=np.float64)
    inner = nu[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(inner, cs * cs * gamma, rtol=1e-12)
    # float32 as the reference computes it: within a few ulp of the same value
    nu32 = les_fd_viscosity(u.astype(np.float32), cs=cs)
    np.testing.assert_allclose(nu32[1:-1, 1:-1, 1:-1], cs * cs * gamma, rtol=2e-6)


def test_rigid_rotation_has_no_strain():
    """u = Omega x r about the z axis: the velocity gradient is antisymmetric,
    S = 0, so nu = 0 (a symmetrisation mistake would leave |S| = sqrt(2) Omega)."""
    x, y, z = _grid()
    om = 0.02
    u = np.zeros((3, N, N, N))
    u[0] = -om * (y - 4.5)
    u[1] = om * (x - 4.5)
    assert np.all(les_fd_viscosity(u, dtype=np.float64) == 0.0)


def test_thresholds_clamp_faces_and_masks():
    x, y, z = _grid()
    u = np.zeros((3, N, N, N))
    # |S| = gamma below the reference's 1e-3 cut: nothing
    u[0] = 5e-4 * z
    assert np.all(les_fd_viscosity(u, dtype=np.float64) == 0.0)
    # zero velocity: nothing (tests/test_les_turbulence.py:135-152)
    assert np.all(les_fd_viscosity(np.zeros_like(u)) == 0.0)
    # a strong shear is clamped to 0.1 (les_turbulence.py:375)
    u[0] = 10.0 * z
    nu = les_fd_viscosity(u, cs=0.18, dtype=np.float64)
    assert np.all(nu[1:-1, 1:-1, 1:-1] == 0.1)
    # faces stay zero
    for sl in (np.s_[0], np.s_[-1], np.s_[:, 0], np.s_[:, -1], np.s_[:, :, 0], np.s_[:, :, -1]):
        assert np.all(nu[sl] == 0.0)
    # the LES mask (filter paper) and the interface suppression |phi| < 0.9
    u[0] = 0.02 * z
    flags = np.zeros((N, N, N), np.uint32)
    flags[4, 4, 4] = L.POL_FLAG_NO_LES
    phase = np.ones((N, N, N))
    phase[5, 5, 5] = 0.5
    phase[6, 6, 6] = -0.95
    nu = les_fd_viscosity(u, mask=mask_from_flags(flags), phase=phase, dtype=np.float64)
    assert nu[4, 4, 4] == 0.0 and nu[5, 5, 5] == 0.0
    assert nu[6, 6, 6] == pytest.approx(0.17 ** 2 * 0.02, rel=1e-12)
    assert nu[3, 3, 3] == pytest.approx(0.17 ** 2 * 0.02, rel=1e-12)


def _tg_state(oracle, n=12, u0=0.04):
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    k = 2 * np.pi / n
    u = np.zeros((3, n, n, n))
    u[0] = u0 * np.sin(k * x) * np.cos(k * y) * np.cos(k * z)
    u[1] = -u0 * np.cos(k * x) * np.sin(k * y) * np.cos(k * z)
    return oracle.init_equilibrium(n, n, n, np.ones((n, n, n)), u)


@pytest.mark.parametrize("tau,nu", [(0.6, 0.05), (0.7, 0.0), (0.52, 0.0), (1.8, 0.1)])
def test_given_viscosity_is_an_effective_tau(oracle, tau, nu):
    """A step with POL_LES | POL_LES_FD and a uniform given nu equals, to FP64
    round-off, a plain step with tau_eff = clamp(tau + 3 nu, 0.55, 1.90)."""
    n = 12
    f = _tg_state(oracle)
    per = oracle.POL_PERIODIC_X | oracle.POL_PERIODIC_Y | oracle.POL_PERIODIC_Z
    d_fd = oracle.make_desc(n, n, n, tau=tau, options=per | oracle.POL_LES | oracle.POL_LES_FD, cs_smag=0.17)
    teff = min(max(tau + 3 * nu, 0.55), 1.90)
    d_plain = oracle.make_desc(n, n, n, tau=teff, options=per)
    a, _, _, _ = oracle.step(d_fd, f, nu=np.full((n, n, n), nu))
    b, _, _, _ = oracle.step(d_plain, f)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-15)
    # without POL_LES_FD the same nu is ignored (the local-stress form runs)
    d_loc = oracle.make_desc(n, n, n, tau=tau, options=per | oracle.POL_LES, cs_smag=0.17)
    c, _, _, _ = oracle.step(d_loc, f, nu=np.full((n, n, n), nu))
    c2, _, _, _ = oracle.step(d_loc, f)
    np.testing.assert_array_equal(c, c2)
