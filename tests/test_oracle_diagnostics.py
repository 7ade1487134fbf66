"""The numpy restatement of the reference's field diagnostics
(oracle/diagnostics.py) pinned by fields whose diagnostics are known in
closed form: uniform states, planted NaN / Inf counts, a linear density
profile, a known momentum change.  CPU only."""
from __future__ import annotations

import math

import numpy as np

from oracle import diagnostics as OD

N3 = (12, 10, 8)


def _uniform(rho0=1.05, u0=(0.03, 0.0, 0.0)):
    rho = np.full(N3, rho0, np.float32)
    u = np.zeros(N3 + (3,), np.float32)
    u[...] = np.array(u0, np.float32)
    return rho, u


def test_umag_operation_order():
    u = np.array([[3e-4, 4e-4, 1.2e-3]], np.float32)
    a, b, c = (np.float32(v) for v in u[0])
    want = np.sqrt(np.float32(np.float32(a * a) + np.float32(b * b)) + np.float32(c * c))
    assert OD.umag(u)[0] == want


def test_uniform_field_quality_and_flow():
    rho, u = _uniform()
    q = OD.lbm_quality(rho, u)
    mach = 0.03 * math.sqrt(3.0)
    assert abs(q["max_mach"] - mach) < 1e-6 and abs(q["mean_mach"] - mach) < 1e-6
    assert q["mach_violation_ratio"] == 0.0
    assert abs(q["max_density_deviation"] - 0.05) < 1e-6
    assert abs(q["mean_density"] - 1.05) < 1e-6 and q["density_variation_coeff"] < 1e-6
    assert q["lbm_grade"] == "Excellent"
    a = OD.flow_analysis(u)
    assert abs(a["max_velocity"] - 0.03) < 1e-8 and abs(a["mean_velocity"] - 0.03) < 1e-8
    assert a["velocity_std"] < 1e-6 and a["flow_complexity"] < 1e-6
    for k in ("top_region_flow", "middle_region_flow", "bottom_region_flow"):
        assert abs(a[k] - 0.03) < 1e-8
    # Mach above 0.1 everywhere: 0.06 * sqrt 3 = 0.1039
    q = OD.lbm_quality(*_uniform(1.0, (0.06, 0.0, 0.0)))
    assert q["mach_violation_ratio"] == 1.0 and q["lbm_grade"] == "Good"


def test_stability_scan_counts_planted_values():
    rho, u = _uniform(1.0, (0.01, 0.0, 0.0))
    solid = np.zeros(N3, np.uint8)
    solid[0] = 1
    rho[3, 3, 3] = np.nan
    rho[4, 4, 4] = np.inf
    rho[0, 2, 2] = np.nan          # solid: not counted
    u[5, 5, 5, 1] = np.nan
    u[6, 6, 6, 2] = -np.inf
    u[7, 7, 7, 0] = 0.4            # max |u| 0.4 > 0.3: warning
    f = np.zeros((19,) + N3, np.float32)
    f[3, 2, 2, 2] = np.nan
    f[7, 2, 2, 2] = np.nan
    f[9, 8, 8, 1] = np.inf
    s = OD.stability_scan(rho, u, solid, f)
    assert (s["nan_rho"], s["inf_rho"], s["nan_u"], s["inf_u"], s["nan_pop"], s["inf_pop"]) == (1, 1, 1, 1, 2, 1)
    assert s["nan_count"] == 4 and s["inf_count"] == 3
    assert s["min_density"] == 1.0 and s["max_density"] == 1.0
    assert abs(s["max_velocity"] - 0.4) < 1e-7
    assert OD.stability_status(s) == 1
    s["nan_count"] = 11
    assert OD.stability_status(s) == 2
    clean = OD.stability_scan(*_uniform(1.0, (0.01, 0, 0)), np.zeros(N3, np.uint8))
    assert OD.stability_status(clean) == 0


def test_stabilize_rules():
    rho, u = _uniform(1.0, (0.01, 0.0, 0.0))
    solid = np.zeros(N3, np.uint8)
    rho[1, 1, 1], rho[2, 2, 2], rho[3, 3, 3], rho[4, 4, 4] = np.nan, -0.5, 7.0, 0.05
    u[5, 5, 5] = (np.nan, 0, 0)
    u[6, 6, 6] = (0.6, 0.8, 0.0)   # |u| = 1 -> scaled to 0.3
    f = np.zeros((19,) + N3, np.float32)
    f[2, 6, 6, 6] = np.nan
    r2, u2, f2 = OD.stabilize(rho, u, solid, f, lambda q, r, uu: np.full(r.shape, 100.0 + q, np.float32))
    assert (r2[1, 1, 1], r2[2, 2, 2], r2[3, 3, 3], r2[4, 4, 4]) == (1.0, 1.0, 5.0, np.float32(0.1))
    assert np.array_equal(u2[5, 5, 5], [0, 0, 0])
    assert abs(np.linalg.norm(u2[6, 6, 6]) - 0.3) < 1e-6
    assert f2[2, 6, 6, 6] == 102.0 and np.isfinite(f2).all()


def test_conservation_and_temporal():
    rho, u = _uniform(1.0, (0.0, 0.0, 0.01))
    st = {"initial_mass": 0.0, "accumulated_inflow": 0.0, "prev_momentum": np.zeros(3)}
    c = OD.conservation(rho, u, st, gravity_lu=1e-5)
    n = np.prod(N3)
    assert c["total_mass"] == n and c["mass_conservation_error"] == 0.0 and c["conservation_grade"] == "Excellent"
    u01 = float(np.float32(0.01))
    assert abs(c["total_momentum"][2] - u01 * n) < 1e-9
    assert abs(c["expected_gravity_change"] - 1e-5 * n) < 1e-12
    # second call: momentum unchanged, mass up 1e-3 relative -> Moderate
    c = OD.conservation(rho * np.float32(1.001), u, st, gravity_lu=0.0)
    assert c["conservation_grade"] == "Moderate"
    assert abs(c["momentum_z_change"] - u01 * n * (float(np.float32(1.001)) - 1.0)) < 1e-9
    assert OD.temporal_stability(rho, u, None, None) == {"status": "first_step"}
    t = OD.temporal_stability(rho, u, rho, u)
    assert t["density_change_rate"] == 0.0 and t["stability_grade"] == "Excellent"


def test_v60_physics_profile_and_outlet():
    rho, u = _uniform(1.0, (0.0, 0.0, -0.02))
    rho[:, :, :5] = 1.2                          # wet below z = 5
    p = OD.v60_physics(rho, u, scale_length=0.5, scale_velocity=2.0)
    assert p["wetting_front_position"] == 4 * 0.5
    nxy = N3[0] * N3[1]
    d12 = float(np.float32(1.2)) - 1.0
    assert abs(p["system_water_volume"] - d12 * nxy * 5 * 0.125) < 1e-9
    assert abs(p["outlet_flow_rate"] - float(np.float32(0.02)) * nxy * 5 * 2.0 * 0.25) < 1e-9


def test_visualizer_interior_statistics():
    rho, u = _uniform(1.0, (0.02, 0.0, 0.0))
    solid = np.zeros(N3, np.uint8)
    phase = np.ones(N3, np.float32)
    phase[:, :, N3[2] // 2:] = 0.0
    s = OD.visualizer_statistics(rho, u, solid, phase)
    inner = (N3[0] - 2) * (N3[1] - 2) * (N3[2] - 2)
    assert s["total_nodes"] == inner
    assert abs(s["total_water_mass"] + s["total_air_mass"] - inner) < 1e-9
    water_z = sum(1 for z in range(1, N3[2] - 1) if z < N3[2] // 2)
    assert s["total_water_mass"] == (N3[0] - 2) * (N3[1] - 2) * water_z
    assert abs(s["max_velocity"] - 0.02) < 1e-8 and abs(s["avg_velocity"] - 0.02) < 1e-8


def test_export_structure():
    rho, u = _uniform(1.0, (0.0, 0.0, 0.0))
    u[1, 2, 3] = (3.0, 4.0, 0.0)
    d = OD.export_data(u, rho, {"step_number": 7})
    assert d["velocity_field"]["magnitude"][1][2][3] == 5.0
    assert abs(d["pressure_field"][0][0][0] - 1.0 / 3.0) < 1e-7
    assert len(d["density_field"]) == N3[0] and len(d["density_field"][0][0]) == N3[2]
