"""numpy restatement of the reference's field diagnostics.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's cpu_baseline leg
as the reference of comparison for the device diagnostics (diagnostics.hip, diagnostics.py)
-- never by the product.

Arrays follow the reference's field order: scalars [nx, ny, nz], vectors
[nx, ny, nz, 3], populations [19, nx, ny, nz] (Taichi field shapes).  Each
function restates the named reference lines.  Element-wise values keep the
reference's float32 arithmetic (|u| = sqrt(ux^2 + uy^2 + uz^2) rounded per
operation, Mach = |u| / float32(1/sqrt 3)); sums and means are taken in
float64 (the reference's float32 numpy accumulations are approximations of
these), std as sqrt(E[x^2] - E[x]^2) of the FP64 sums.
"""
from __future__ import annotations

import json
from typing import Optional

import numpy as np

CS_F32 = np.float32(1.0 / np.sqrt(3.0))


def umag(u: np.ndarray) -> np.ndarray:
    """float32 |u| in the reference's operation order (lbm_diagnostics.py:200)."""
    u = u.astype(np.float32, copy=False)
    return np.sqrt((u[..., 0] * u[..., 0] + u[..., 1] * u[..., 1]) + u[..., 2] * u[..., 2])


def _std(x: np.ndarray) -> float:
    x = x.astype(np.float64).ravel()
    m = x.mean()
    return float(np.sqrt(max(np.mean(x * x) - m * m, 0.0)))


# ---- numerical_stability.py:53-110 -------------------------------------------------
def stability_scan(rho: np.ndarray, u: np.ndarray, solid: np.ndarray, f: Optional[np.ndarray] = None) -> dict:
    fluid = solid == 0
    r = rho[fluid].astype(np.float64)
    um = umag(u)[fluid].astype(np.float64)
    out = {"nan_rho": int(np.isnan(r).sum()), "inf_rho": int(np.isinf(r).sum())}
    fin = np.isfinite(r)
    out["min_density"] = float(r[fin].min()) if fin.any() else 1e10
    out["max_density"] = float(r[fin].max()) if fin.any() else 0.0
    out["nan_u"] = int(np.isnan(um).sum())
    out["inf_u"] = int(np.isinf(um).sum())
    finu = np.isfinite(um)
    out["max_velocity"] = float(um[finu].max()) if finu.any() else 0.0
    out["nan_pop"] = out["inf_pop"] = 0
    if f is not None:
        ff = f[:, fluid].astype(np.float64)
        out["nan_pop"] = int(np.isnan(ff).sum())
        out["inf_pop"] = int(np.isinf(ff).sum())
    out["nan_count"] = out["nan_rho"] + out["nan_u"] + out["nan_pop"]
    out["inf_count"] = out["inf_rho"] + out["inf_u"] + out["inf_pop"]
    return out


def stability_status(scan: dict, max_v: float = 0.3, min_rho: float = 0.1, max_rho: float = 5.0,
                     max_nan: int = 10) -> int:
    """numerical_stability.py:99-110: 0 normal, 1 warning, 2 severe."""
    if (scan["nan_count"] > 0 or scan["inf_count"] > 0 or scan["max_velocity"] > max_v
            or scan["min_density"] < min_rho or scan["max_density"] > max_rho):
        if scan["nan_count"] > max_nan or scan["max_velocity"] > 0.5:
            return 2
        return 1
    return 0


# ---- numerical_stability.py:165-197 -------------------------------------------------
def stabilize(rho: np.ndarray, u: np.ndarray, solid: np.ndarray, f: np.ndarray, feq, max_v=0.3, min_rho=0.1,
              max_rho=5.0):
    """Returns repaired copies (rho, u, f); feq(q, rho, u) -> population value.
    f holds the stored deviations; a non-finite one becomes feq of the
    repaired (rho, u)."""
    rho, u, f = rho.copy(), u.copy(), f.copy()
    fluid = solid == 0
    r = rho[fluid]
    bad = ~np.isfinite(r) | (r <= 0)
    r = np.where(bad, np.float32(1), np.where(r > max_rho, np.float32(max_rho),
                                              np.where(r < min_rho, np.float32(min_rho), r)))
    rho[fluid] = r
    uu = u[fluid]
    m = umag(uu)
    badu = ~np.isfinite(m)
    scale = np.where(badu | (m <= max_v), np.float32(1), (np.float32(max_v) / np.where(m > 0, m, 1)).astype(np.float32))
    uu = (uu * scale[:, None]).astype(u.dtype)
    uu[badu] = 0
    u[fluid] = uu
    idx = np.nonzero(fluid)
    for q in range(f.shape[0]):
        fq = f[q][idx]
        nb = ~np.isfinite(fq)
        if nb.any():
            fq[nb] = feq(q, r[nb], uu[nb])
            f[q][idx] = fq
    return rho, u, f


# ---- lbm_diagnostics.py:151-186 ----------------------------------------------------------
def temporal_stability(rho, u, prev_rho, prev_u, dt=1.0) -> dict:
    if prev_rho is None or prev_u is None:
        return {"status": "first_step"}
    dr = np.sqrt(np.sum((rho.astype(np.float64) - prev_rho) ** 2))
    du = np.sqrt(np.sum((u.astype(np.float64) - prev_u) ** 2))
    out = {"density_change_rate": dr / dt, "relative_density_change": dr / np.sqrt(np.sum(rho.astype(np.float64) ** 2)),
           "velocity_change_rate": du / dt,
           "relative_velocity_change": du / (np.sqrt(np.sum(u.astype(np.float64) ** 2)) + 1e-10)}
    if out["relative_density_change"] < 1e-6 and out["relative_velocity_change"] < 1e-4:
        out["stability_grade"] = "Excellent"
    elif out["relative_density_change"] < 1e-4 and out["relative_velocity_change"] < 1e-2:
        out["stability_grade"] = "Good"
    else:
        out["stability_grade"] = "Monitoring"
    return out


# ---- lbm_diagnostics.py:188-226 ------------------------------------------------------------
def lbm_quality(rho, u, reference_density=1.0) -> dict:
    mach = umag(u) / CS_F32
    q = {"max_mach": float(np.max(mach))}
    act = mach[mach > np.float32(1e-10)]
    q["mean_mach"] = float(np.mean(act.astype(np.float64))) if act.size else 0.0
    q["mach_violation_ratio"] = float(np.sum(mach > np.float32(0.1)) / mach.size)
    dev = np.abs(rho.astype(np.float64) - reference_density)
    q["max_density_deviation"] = float(np.max(dev))
    q["mean_density"] = float(np.mean(rho.astype(np.float64)))
    q["density_variation_coeff"] = _std(rho) / q["mean_density"]
    if q["max_mach"] < 0.1 and q["max_density_deviation"] < 0.1:
        q["lbm_grade"] = "Excellent"
    elif q["max_mach"] < 0.2 and q["max_density_deviation"] < 0.2:
        q["lbm_grade"] = "Good"
    else:
        q["lbm_grade"] = "Caution"
    return q


# ---- lbm_diagnostics.py:228-291 ------------------------------------------------------------
def conservation(rho, u, state: dict, gravity_lu: float, dt: float = 1.0, inflow_rate: float = 0.0) -> dict:
    """state: {'initial_mass', 'accumulated_inflow', 'prev_momentum'} (updated)."""
    mass = float(np.sum(rho.astype(np.float64)))
    if state["initial_mass"] == 0.0:
        state["initial_mass"] = mass
    state["accumulated_inflow"] += inflow_rate * dt
    expected = state["initial_mass"] + state["accumulated_inflow"]
    err = abs(mass - expected)
    c = {"total_mass": mass, "initial_mass": state["initial_mass"], "accumulated_inflow": state["accumulated_inflow"],
         "mass_conservation_error": err, "relative_mass_error": err / state["initial_mass"]}
    mom = np.sum(rho.astype(np.float64)[..., None] * u.astype(np.float64), axis=(0, 1, 2))
    change = mom - state["prev_momentum"]
    g = gravity_lu * mass * dt
    c["total_momentum"] = mom.tolist()
    c["momentum_z_change"] = float(change[2])
    c["expected_gravity_change"] = float(g)
    c["momentum_conservation_error"] = float(abs(change[2] - g))
    c["conservation_grade"] = ("Excellent" if c["relative_mass_error"] < 1e-6 else
                               "Good" if c["relative_mass_error"] < 1e-4 else "Moderate")
    state["prev_momentum"] = mom.copy()
    return c


# ---- lbm_diagnostics.py:328-370, 445-460 ---------------------------------------------------
def outlet_flow(u, scale_velocity: float, scale_length: float) -> float:
    bottom = u[:, :, 0:5, 2].astype(np.float64)
    return float(np.sum(np.maximum(-bottom, 0)) * scale_velocity * scale_length ** 2)


def v60_physics(rho, u, scale_length: float, scale_velocity: float, reference_density=1.0,
                inlet_flow_rate=0.0) -> dict:
    r = rho.astype(np.float64)
    prof = np.mean(r, axis=(0, 1))
    wet = np.where(prof > reference_density + 0.1)[0]
    p = {"wetting_front_position": float(np.max(wet) * scale_length) if wet.size else 0.0,
         "system_water_volume": float(max(0.0, np.sum(r - reference_density) * scale_length ** 3)),
         "inlet_flow_rate": float(inlet_flow_rate),
         "outlet_flow_rate": outlet_flow(u, scale_velocity, scale_length)}
    return p


# ---- lbm_diagnostics.py:372-407 --------------------------------------------------------------
def flow_analysis(u) -> dict:
    m32 = umag(u)
    m = m32.astype(np.float64)
    a = {"max_velocity": float(np.max(m))}
    act = m[m32 > np.float32(1e-10)]
    a["mean_velocity"] = float(np.mean(act)) if act.size else 0.0
    a["velocity_std"] = _std(m)
    a["flow_complexity"] = float(np.linalg.norm([_std(u[..., c]) for c in range(3)]))
    nz = u.shape[2]
    a["top_region_flow"] = float(np.mean(m[:, :, int(nz * 0.8):]))
    a["middle_region_flow"] = float(np.mean(m[:, :, int(nz * 0.3):int(nz * 0.8)]))
    a["bottom_region_flow"] = float(np.mean(m[:, :, :int(nz * 0.3)]))
    return a


# ---- visualizer.py:120-167 ---------------------------------------------------------------------
def visualizer_statistics(rho, u, solid, phase) -> dict:
    nx, ny, nz = rho.shape
    inner = np.zeros(rho.shape, bool)
    inner[1:nx - 1, 1:ny - 1, 1:nz - 1] = True
    sel = inner & (solid == 0)
    r = rho[sel].astype(np.float64)
    m = umag(u)[sel].astype(np.float64)
    water = phase[sel] > 0.1
    n = int(sel.sum())
    return {"total_water_mass": float(np.sum(r[water])), "total_air_mass": float(np.sum(r[~water])),
            "max_velocity": float(m.max()) if n else 0.0, "min_velocity": float(m.min()) if n else 0.0,
            "avg_velocity": float(m.sum() / n) if n else 0.0, "total_nodes": n}


# ---- enhanced_visualizer.py:1487-1559 ----------------------------------------------------------
def export_data(u, rho, metadata: dict, particles: Optional[dict] = None, cs2: float = 1.0 / 3.0) -> dict:
    """The exported dictionary (u [nx, ny, nz, 3], rho [nx, ny, nz] float32)."""
    d = {"velocity_field": {"u_x": u[..., 0].tolist(), "u_y": u[..., 1].tolist(), "u_z": u[..., 2].tolist(),
                            "magnitude": np.sqrt(np.sum(u ** 2, axis=3)).tolist()},
         "density_field": rho.tolist(), "pressure_field": (rho * cs2).tolist()}
    if particles is not None:
        act = particles["active"] == 1
        d["particles"] = {"positions": particles["positions"][act].tolist(),
                          "velocities": particles["velocities"][act].tolist(), "count": int(np.sum(particles["active"]))}
    d["metadata"] = metadata
    return d


def export_json_text(d: dict) -> str:
    return json.dumps(d, indent=2)
