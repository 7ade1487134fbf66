"""numpy restatement of the reference's thermal D3Q7 solver.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's cpu_baseline leg
as the reference of comparison for csrc/thermal.hip -- never by the product.

Follows src/physics/thermal_lbm.py in the reference's array order (fields
[nx, ny, nz], populations [nx, ny, nz, 7], q = rest, +x, -x, +y, -y, +z, -z)
and its expressions operation by operation, in the populations' dtype
(float32 as in the reference, or float64; numpy performs no contraction):
  T = ((((((0 + g0) + g1) + ...) + g6)            thermal_lbm.py:214-218
  g_new = (g - omega (g - w T)) + (w S) DT         :220-229
  g(x) = g_new(x - e) inside the box, else g_new(x) :231-253
  S += -((ux dT/dx + uy dT/dy) + uz dT/dz), dT/dx = (T[x+1] - T[x-1]) / (2 DX)
                                                     :518-538 (interior cells)
"""
from __future__ import annotations

import numpy as np

CX = np.array([0, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0])
CZ = np.array([0, 0, 0, 0, 0, 1, -1])
W = np.array([0.25] + [0.125] * 6, dtype=np.float32)
DX = np.float32(1.0)
DT = np.float32(1.0)


def equilibrium(T: np.ndarray) -> np.ndarray:
    """thermal_lbm.py:186-201 for every q: [nx, ny, nz, 7]."""
    T = np.asarray(T)
    if T.dtype != np.float64:
        T = T.astype(np.float32)
    return np.stack([W[q] * T for q in range(7)], axis=-1)


def temperature(g: np.ndarray) -> np.ndarray:
    """thermal_lbm.py:255-266."""
    T = np.zeros(g.shape[:3], g.dtype)
    for q in range(7):
        T = T + g[..., q]
    return T


def convection_source(T: np.ndarray, u: np.ndarray, hs: np.ndarray) -> np.ndarray:
    """thermal_lbm.py:518-538: hs + (-u.grad T) on interior cells; u [nx, ny, nz, 3]."""
    hs = hs.copy()
    two = np.float32(2.0) * DX
    c = (slice(1, -1),) * 3
    dTdx = (T[2:, 1:-1, 1:-1] - T[:-2, 1:-1, 1:-1]) / two
    dTdy = (T[1:-1, 2:, 1:-1] - T[1:-1, :-2, 1:-1]) / two
    dTdz = (T[1:-1, 1:-1, 2:] - T[1:-1, 1:-1, :-2]) / two
    uu = u[c]
    conv = -((uu[..., 0] * dTdx + uu[..., 1] * dTdy) + uu[..., 2] * dTdz)
    hs[c] = hs[c] + conv
    return hs


def collide(g: np.ndarray, hs: np.ndarray, omega: np.float32):
    """thermal_lbm.py:203-229: (T_local, g_new)."""
    T = temperature(g)
    gn = np.empty_like(g)
    for q in range(7):
        geq = W[q] * T
        src = (W[q] * hs) * DT
        gn[..., q] = (g[..., q] - omega * (g[..., q] - geq)) + src
    return T, gn


def stream(gn: np.ndarray) -> np.ndarray:
    """thermal_lbm.py:231-253 (an out-of-box source keeps the own value)."""
    nx, ny, nz = gn.shape[:3]
    g = gn.copy()
    for q in range(1, 7):
        sx = slice(max(CX[q], 0), nx + min(CX[q], 0))
        sy = slice(max(CY[q], 0), ny + min(CY[q], 0))
        sz = slice(max(CZ[q], 0), nz + min(CZ[q], 0))
        tx = slice(max(-CX[q], 0), nx + min(-CX[q], 0))
        ty = slice(max(-CY[q], 0), ny + min(-CY[q], 0))
        tz = slice(max(-CZ[q], 0), nz + min(-CZ[q], 0))
        g[sx, sy, sz, q] = gn[tx, ty, tz, q]
    return g


def step(state: dict, omega: float, convection: bool = False) -> dict:
    """thermal_lbm.py:396-427 (without the flux pass): state holds g, T,
    heat_source (and u [nx, ny, nz, 3] with convection); returns the new state."""
    om = state["g"].dtype.type(omega)
    s = dict(state)
    s["T_old"] = state["T"]
    if convection:
        s["heat_source"] = convection_source(state["T"], state["u"], state["heat_source"])
    _, gn = collide(state["g"], s["heat_source"], om)
    s["g"] = stream(gn)
    s["T"] = temperature(s["g"])
    return s


def heat_flux(T: np.ndarray, k: np.ndarray) -> np.ndarray:
    """thermal_lbm.py:268-286: [nx, ny, nz, 3], zero on the faces."""
    q = np.zeros(T.shape + (3,), T.dtype)
    two = np.float32(2.0) * DX
    c = (slice(1, -1),) * 3
    kk = k[c].astype(T.dtype)
    q[c + (0,)] = -kk * ((T[2:, 1:-1, 1:-1] - T[:-2, 1:-1, 1:-1]) / two)
    q[c + (1,)] = -kk * ((T[1:-1, 2:, 1:-1] - T[1:-1, :-2, 1:-1]) / two)
    q[c + (2,)] = -kk * ((T[1:-1, 1:-1, 2:] - T[1:-1, 1:-1, :-2]) / two)
    return q


def unstable_cells(T: np.ndarray, g: np.ndarray) -> int:
    """thermal_lbm.py:348-376, counted per cell."""
    bad = (T < -50.0) | (T > 150.0) | np.isnan(T)
    bad |= np.any(np.isnan(g) | (np.abs(g) > 1000.0), axis=-1)
    return int(bad.sum())


def boundary(state: dict, kind: int, mask: np.ndarray, a: float, b: float = 0.0) -> dict:
    """thermal_lbm.py:288-346: 1 Dirichlet, 2 Neumann, 3 convective."""
    s = dict(state)
    ty = state["T"].dtype.type
    if kind == 1:
        T = state["T"].copy()
        g = state["g"].copy()
        T[mask] = ty(a)
        g[mask] = W * ty(a)
        s["T"], s["g"] = T, g
    elif kind == 2:
        hs = state["heat_source"].copy()
        hs[mask] = ty(a)
        s["heat_source"] = hs
    else:
        hs = state["heat_source"].copy()
        hs[mask] = -(ty(a) * (state["T"][mask] - ty(b)))
        s["heat_source"] = hs
    return s
