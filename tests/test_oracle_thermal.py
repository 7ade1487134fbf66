"""The numpy restatement of the reference's thermal D3Q7 solver
(oracle/thermal.py) pinned by closed forms: uniform invariance, the one-step
response to a delta, conservation away from the faces, the diffusion rate of
a Gaussian, exact convection source and heat flux of linear fields, the
boundary forms and the stability test.  CPU only."""
from __future__ import annotations

import numpy as np

from oracle import thermal as OT


def _state(T, hs=None, u=None):
    T = T.astype(np.float32)
    s = {"T": T, "g": OT.equilibrium(T), "heat_source": np.zeros_like(T) if hs is None else hs.astype(np.float32)}
    if u is not None:
        s["u"] = u.astype(np.float32)
    return s


def test_weights_and_uniform_invariance():
    assert OT.W.sum() == 1.0
    assert np.array_equal(OT.W[1:] @ np.stack([OT.CX[1:] ** 2, OT.CY[1:] ** 2, OT.CZ[1:] ** 2], 1), [0.25] * 3)
    s = _state(np.full((6, 5, 4), 80.0))
    for _ in range(5):
        s = OT.step(s, omega=1 / 0.7)
    assert np.array_equal(s["T"], np.full((6, 5, 4), 80.0, np.float32))


def test_delta_one_step():
    T = np.zeros((7, 7, 7))
    T[3, 3, 3] = 1.0
    s = OT.step(_state(T), omega=1.3)
    want = np.zeros((7, 7, 7), np.float32)
    want[3, 3, 3] = 0.25
    for dx, dy, dz in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        want[3 + dx, 3 + dy, 3 + dz] = 0.125
    assert np.array_equal(s["T"], want)


def test_conservation_and_gaussian_spreading():
    n, tau, steps = 48, 0.8, 60
    x = np.arange(n) - (n - 1) / 2
    r2 = x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2
    T = np.exp(-r2 / (2 * 3.0 ** 2))
    s = _state(T)
    tot0 = float(s["T"].astype(np.float64).sum())
    var0 = float((s["T"] * x[:, None, None] ** 2).sum() / s["T"].sum())
    for _ in range(steps):
        s = OT.step(s, omega=1 / tau)
    Tn = s["T"].astype(np.float64)
    assert abs(Tn.sum() - tot0) < 1e-5 * tot0          # the blob never reaches a face
    var = float((Tn * x[:, None, None] ** 2).sum() / Tn.sum())
    D = 0.25 * (tau - 0.5)                              # D3Q7: c_s^2 = sum w e_x^2 = 1/4
    assert abs((var - var0) / (2 * D * steps) - 1) < 0.02


def test_convection_source_and_heat_flux_of_linear_fields():
    nx, ny, nz = 6, 5, 7
    X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    T = (2.0 * X - 3.0 * Y + 0.5 * Z).astype(np.float32)
    u = np.zeros((nx, ny, nz, 3), np.float32)
    u[...] = (0.1, 0.2, -0.4)
    hs = np.full((nx, ny, nz), 1.0, np.float32)
    out = OT.convection_source(T, u, hs)
    want = 1.0 - (0.1 * 2.0 + 0.2 * -3.0 + -0.4 * 0.5)
    assert np.allclose(out[1:-1, 1:-1, 1:-1], want, rtol=0, atol=1e-6)
    assert np.array_equal(out[0], hs[0]) and np.array_equal(out[:, :, -1], hs[:, :, -1])
    q = OT.heat_flux(T, np.full((nx, ny, nz), 0.5, np.float32))
    assert np.allclose(q[1:-1, 1:-1, 1:-1], (-1.0, 1.5, -0.25), atol=0)
    assert not q[0].any() and not q[:, -1].any()


def test_boundary_forms_and_stability():
    T = np.full((5, 5, 5), 20.0, np.float32)
    s = _state(T, hs=np.full((5, 5, 5), 0.5))
    m = np.zeros((5, 5, 5), bool)
    m[0] = True
    d = OT.boundary(s, 1, m, 95.0)
    assert (d["T"][0] == 95.0).all() and np.array_equal(d["g"][0, 0, 0], OT.W * np.float32(95.0))
    assert (d["T"][1:] == 20.0).all()
    n = OT.boundary(s, 2, m, 3.0)
    assert (n["heat_source"][0] == 3.0).all() and (n["heat_source"][1:] == 0.5).all()
    c = OT.boundary(s, 3, m, 2.0, 25.0)
    assert (c["heat_source"][0] == 10.0).all()
    assert OT.unstable_cells(s["T"], s["g"]) == 0
    g = s["g"].copy()
    g[1, 1, 1, 3] = np.nan
    g[2, 2, 2, 0] = 2000.0
    T2 = s["T"].copy()
    T2[3, 3, 3] = 151.0
    T2[4, 4, 4] = -60.0
    assert OT.unstable_cells(T2, g) == 4
