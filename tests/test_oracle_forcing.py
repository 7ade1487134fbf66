"""Pins of the body-force oracle (oracle/forcing.py) by closed forms: the
pour region's shape and intensities, the acceleration and phase updates,
the target-density profile, the pressure force of a linear density, the
density-drive limiter, the blockage law and the particle bounce."""
from __future__ import annotations

import math

import numpy as np

from oracle import forcing as OF

F = np.float32
SHAPE = (24, 20, 22)  # nz, ny, nx


def test_pour_region_shape_and_intensity():
    tot = OF.pour_intensity(SHAPE, 10.0, 9.0, 4.0, 20)
    assert tot[20, 9, 10] == 1.0                           # nozzle centre: exp(0) exp(0)
    np.testing.assert_allclose(tot[16, 9, 10], math.exp(-2.0), rtol=1e-6)
    np.testing.assert_allclose(tot[20, 9, 14], math.exp(-0.5), rtol=1e-6)  # on the rim
    assert np.count_nonzero(tot) == 49 * 5                # |d| <= 4 disc (49 cells) x 5 layers
    assert not tot[15].any() and not tot[21].any()


def test_pour_force_and_phase_closed_forms():
    solid = np.zeros(SHAPE, np.uint8)
    solid[20, 9, 11] = 1
    bf = np.zeros((3,) + SHAPE, F)
    bf[0, 20, 9, 10] = F(-0.0)
    out = OF.pour_force(bf, solid, 10.0, 9.0, 4.0, 20, 0.05, 1.0, 1.0)
    assert out[2, 20, 9, 10] == F(-0.05)
    assert np.signbit(out[0, 20, 9, 10]) == False          # -0 + 0 = +0
    assert out[2, 20, 9, 11] == 0                          # solid cell untouched
    assert (out[2] < 0).sum() == 49 * 5 - 1
    big = OF.pour_force(bf, solid, 10.0, 9.0, 4.0, 20, 0.05, 1.0, 1e-4)
    assert big[2, 20, 9, 10] == F(-10.0)                   # clamp at 10
    none = OF.pour_force(bf, solid, 10.0, 9.0, 4.0, 20, 0.05, 1.0, 1e-9)
    assert not none[2].any()                               # dt <= 1e-8: no acceleration
    phi = np.full(SHAPE, F(-1.0))
    p1 = OF.pour_phase(phi, solid, 10.0, 9.0, 4.0, 20, 1.0, 0.1)
    np.testing.assert_allclose(p1[20, 9, 10], -0.8, rtol=1e-6)   # rate clamped to 2
    assert OF.pour_phase(phi, solid, 10.0, 9.0, 4.0, 20, 1.0, 1.0)[20, 9, 10] == 1.0


def test_spiral_position():
    x, y = OF.spiral_position(0.0, 1.0, 5.0, 100.0, 90.0, 8.0, 224, 224)
    assert (x, y) == (F(105.0), F(90.0))
    x, y = OF.spiral_position(math.pi / 2, 1.0, 5.0, 100.0, 90.0, 8.0, 224, 224)
    r = 5.0 * (1 + 0.1 * math.pi / 2)
    np.testing.assert_allclose([x, y], [100.0, 90.0 + r], atol=2e-5)
    x, _ = OF.spiral_position(0.0, 1.0, 500.0, 100.0, 90.0, 8.0, 224, 224)
    assert x == F(216.0)                                   # clamped to NX - diameter


def test_target_density_profile():
    t = OF.target_density((100, 2, 2))
    assert t[0, 0, 0] == F(0.4)
    np.testing.assert_allclose(t[10, 0, 0], 0.4 + 0.5 * 0.6, rtol=1e-6)
    assert t[50, 0, 0] == 1.0
    np.testing.assert_allclose(t[90, 0, 0], 1.0 + 0.5 * 0.8, rtol=1e-6)


def test_pressure_force_of_linear_density():
    k = np.arange(SHAPE[0], dtype=F)
    rho = np.broadcast_to((F(1.0) + F(0.0078125) * k)[:, None, None], SHAPE).astype(F)
    solid = np.zeros(SHAPE, np.uint8)
    pf = OF.pressure_force(rho, solid)
    want = -(0.0078125 / 3.0) / rho.astype(np.float64)
    np.testing.assert_allclose(pf[2], want, rtol=1e-6)     # faces one-sided: same slope
    assert not pf[0].any() and not pf[1].any()
    steep = np.broadcast_to((F(1.0) + F(0.5) * k)[:, None, None], SHAPE).astype(F)
    m = np.sqrt((OF.pressure_force(steep, solid).astype(np.float64) ** 2).sum(axis=0))
    np.testing.assert_allclose(m, np.minimum(0.12, (0.5 / 3.0) / steep), rtol=1e-6)  # norm limit
    assert (m == F(0.12)).any() or np.isclose(m, 0.12, rtol=1e-7).any()
    solid[3] = 1
    pf2 = OF.pressure_force(rho, solid, pf_prev=np.full((3,) + SHAPE, F(7)))
    assert np.all(pf2[:, 3] == 7)
    bf = OF.accumulate(np.ones((3,) + SHAPE, F), pf2, solid, half=True)
    assert np.all(bf[:, 3] == 1)
    np.testing.assert_array_equal(bf[2, 5], F(1) + F(0.5) * pf2[2, 5])


def test_density_drive_limiter():
    solid = np.zeros(SHAPE, np.uint8)
    rho = np.ones(SHAPE, F)
    tgt = np.full(SHAPE, F(1.8))
    assert np.all(OF.density_drive(rho, tgt, solid) == F(1.001))
    near = np.full(SHAPE, F(1.02))
    np.testing.assert_allclose(OF.density_drive(rho, near, solid), 1.0005, rtol=1e-6)
    assert np.all(OF.density_drive(np.full(SHAPE, F(2.5)), tgt, solid) == F(2.0))
    assert np.all(OF.density_drive(rho, tgt, solid, half=True, max_adj=0.0005, lo=0.7, hi=1.5) == F(1.0005))


def test_filter_blockage_law():
    zone = np.zeros(SHAPE, np.uint8)
    zone[2] = 1
    acc = np.full(SHAPE, F(10.0))
    blk = np.full(SHAPE, F(0.2))
    a, b = OF.filter_blockage(zone, acc, blk)
    np.testing.assert_allclose(b[2], 0.95 * 0.2 + 0.05 * 0.9 * (1 - math.exp(-1.0)), rtol=1e-6)
    np.testing.assert_allclose(a[2], 9.99, rtol=1e-6)
    assert np.all(b[3] == F(0.2)) and np.all(a[3] == F(10))


def test_block_particles_bounce():
    zone = np.zeros(SHAPE, np.uint8)
    zone[5] = 1
    acc = np.zeros(SHAPE, F)
    s = 0.5
    pos = np.array([[3.2, 4.1, 6.3], [5.2, 6.1, 3.3], [7.2, 8.1, 3.3], [1.0, 1.0, 20.0]], F).T * F(s)
    vel = np.array([[0.1, 0.0, -0.2], [0.0, 0.1, 0.3], [0.0, 0.0, -0.1], [0, 0, -1]], F).T
    active = np.array([1, 1, 0, 1], np.int32)
    a, v = OF.block_particles(zone, acc, pos, vel, active, s, seed=7)
    np.testing.assert_allclose(v[2, 0], 0.06, rtol=1e-6)    # -(-0.2) x 0.3
    assert abs(v[0, 0] - 0.1) <= 0.005 and abs(v[1, 0]) <= 0.005
    assert a[5, 4, 3] == F(0.01) and a.sum() == F(0.01)
    np.testing.assert_array_equal(v[:, 1:], vel[:, 1:])       # moving up / inactive / far from the filter
