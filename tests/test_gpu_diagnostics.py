"""Device field diagnostics (diagnostics.hip through pol_field_stats /
pol_stabilize / pol_export_fields, and the reference-mirroring classes of
pourover_amd.diagnostics) against the numpy restatement of the reference's
diagnostics (oracle/diagnostics.py) on the same fields.

Counts and extremes compare exactly (|u| is formed per element in the same
float32 operations); FP64 sums and the values derived from them compare to
rtol 1e-10 (block-order FP64 sums against numpy's pairwise sums).  The export
is compared byte for byte as JSON text."""
from __future__ import annotations

import json

import numpy as np
import pytest
import torch

from oracle import diagnostics as OD
from pourover_amd import _lib as L
from pourover_amd import diagnostics as D
from pourover_amd import workloads as WL
from pourover_amd.engine import HipContext

pytestmark = pytest.mark.gpu

RT = 1e-10


def _xyz(t):
    return np.ascontiguousarray(t.detach().cpu().numpy().transpose(2, 1, 0))


def _uxyz(t):
    return np.ascontiguousarray(t.detach().cpu().numpy().transpose(3, 2, 1, 0))


def _close(a, b, rt=RT, at=1e-13):
    return abs(a - b) <= rt * max(abs(a), abs(b)) + at


def _state(precision="f32", n=48, steps=150):
    wl = WL.v60_config(n, precision)
    ctx = HipContext(wl.config, wl.geometry.flags)
    ctx.init_equilibrium()
    ctx.step(steps)
    ctx.macroscopic(macro=True)
    torch.cuda.synchronize()
    solid = (wl.geometry.kind == L.POL_KIND_SOLID).transpose(2, 1, 0).astype(np.uint8)
    return wl, ctx, solid


@pytest.mark.parametrize("precision", ["f32", "f16", "f64"])
def test_field_statistics_match_oracle(precision):
    print("case: field statistics V60 48^3 %s" % precision, flush=True)
    wl, ctx, solid = _state(precision)
    rho, u = _xyz(ctx.rho), _uxyz(ctx.u)
    prev_rho = ctx.rho * 0.999 + 0.001
    prev_u = ctx.u * 0.9
    phase = torch.rand(ctx.rho.shape, generator=torch.Generator().manual_seed(5)).to(ctx.rho)
    s, zp = D.field_statistics(ctx, prev_rho, prev_u, phase, populations=True)
    n = rho.size
    r64, m = rho.astype(np.float64), OD.umag(u).astype(np.float64)
    assert s["cells"] == n
    assert _close(s["sum_drho"], np.sum(r64 - 1.0)) and _close(s["sum_drho2"], np.sum((r64 - 1.0) ** 2))
    assert s["max_abs_drho"] == np.max(np.abs(r64 - 1.0))
    mom = np.sum(r64[..., None] * u.astype(np.float64), axis=(0, 1, 2))
    for i, k in enumerate(("mom_x", "mom_y", "mom_z")):
        assert _close(s[k], mom[i], at=1e-12 * n)
    assert _close(s["sum_umag"], m.sum()) and _close(s["sum_umag2"], (m * m).sum())
    assert s["max_umag"] == m.max()
    mach = OD.umag(u) / OD.CS_F32
    assert s["active"] == np.sum(mach > np.float32(1e-10)) and s["mach_gt"] == np.sum(mach > np.float32(0.1))
    assert _close(s["sum_mach_active"], mach[mach > np.float32(1e-10)].astype(np.float64).sum())
    assert s["umag_active"] == np.sum(m > np.float32(1e-10))
    nz = rho.shape[2]
    zt, zm = int(nz * 0.8), int(nz * 0.3)
    assert _close(s["sum_umag_top"], m[:, :, zt:].sum()) and _close(s["sum_umag_mid"], m[:, :, zm:zt].sum())
    assert _close(s["sum_umag_bot"], m[:, :, :zm].sum())
    assert _close(s["outlet_down"], np.maximum(-u[:, :, 0:5, 2].astype(np.float64), 0).sum())
    assert _close(s["drho_prev2"], np.sum((r64 - _xyz(prev_rho).astype(np.float64)) ** 2))
    assert _close(s["du_prev2"], np.sum((u.astype(np.float64) - _uxyz(prev_u).astype(np.float64)) ** 2))
    assert np.allclose(zp, r64.sum(axis=(0, 1)), rtol=RT, atol=0)
    f = np.ascontiguousarray(ctx.dense_populations().cpu().numpy().transpose(0, 3, 2, 1))
    scan = OD.stability_scan(rho, u, solid, f)
    assert s["fluid"] == np.sum(solid == 0)
    assert (s["fluid_rho_nan"], s["fluid_rho_inf"], s["fluid_u_nan"], s["fluid_u_inf"]) == (0, 0, 0, 0)
    assert (s["fluid_pop_nan"], s["fluid_pop_inf"]) == (0, 0)
    assert s["fluid_rho_min"] == scan["min_density"] and s["fluid_rho_max"] == scan["max_density"]
    assert s["fluid_umax"] == scan["max_velocity"]
    vis = OD.visualizer_statistics(rho, u, solid, _xyz(phase))
    assert s["int_cells"] == vis["total_nodes"]
    assert _close(s["int_water"], vis["total_water_mass"]) and _close(s["int_air"], vis["total_air_mass"])
    assert s["int_umax"] == vis["max_velocity"] and s["int_umin"] == vis["min_velocity"]
    assert _close(s["int_usum"] / s["int_cells"], vis["avg_velocity"])


def test_planted_nonfinite_values_and_recovery():
    print("case: planted NaN / Inf, stability monitor and recovery, V60 48^3 f32", flush=True)
    wl, ctx, solid = _state("f32")
    fl = np.argwhere(solid.transpose(2, 1, 0) == 0)  # (z, y, x) of non-solid cells
    rng = np.random.default_rng(11)
    pick = fl[rng.choice(len(fl), 12, replace=False)]
    (z0, y0, x0), (z1, y1, x1), (z2, y2, x2), (z3, y3, x3) = pick[:4]
    ctx.rho[z0, y0, x0] = float("nan")
    ctx.rho[z1, y1, x1] = float("inf")
    ctx.u[1, z2, y2, x2] = float("nan")
    ctx.u[:, z3, y3, x3] = torch.tensor([0.6, 0.8, 0.0])   # |u| = 1: severe (> 0.5)
    f = ctx.f_current
    for k, (z, y, x) in enumerate(pick[4:10]):
        slot = int(ctx.layout.slot_of(x, y, z)) if ctx.layout is not None else int((z * ctx.config.ny + y) * ctx.config.nx + x)
        f[3 + k, slot] = float("nan") if k < 4 else float("inf")
    torch.cuda.synchronize()
    rho, u = _xyz(ctx.rho), _uxyz(ctx.u)
    fd = np.ascontiguousarray(ctx.dense_populations().cpu().numpy().transpose(0, 3, 2, 1))
    scan = OD.stability_scan(rho, u, solid, fd)
    assert (scan["nan_pop"], scan["inf_pop"]) == (4, 2)
    mon = D.NumericalStabilityMonitor()
    rep = mon.diagnose_stability(ctx, step=7)
    assert rep["status"] == OD.stability_status(scan) == 2
    assert (rep["nan_count"], rep["inf_count"]) == (scan["nan_count"], scan["inf_count"])
    assert rep["max_velocity"] == scan["max_velocity"]
    assert rep["min_density"] == scan["min_density"] and rep["max_density"] == scan["max_density"]
    assert mon.consecutive_errors == 1 and not mon.check_stability(ctx)
    # emergency stabilisation against the restated rules

    def feq(q, r, uu):
        from pourover_amd.lattice import E
        w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
        r = r.astype(np.float64)
        uu = uu.astype(np.float64)
        eu = uu @ E[q].astype(np.float64)
        usq = (uu * uu).sum(axis=-1)
        return (36 * w[q] * ((r - 1) + r * (3 * eu + 4.5 * eu * eu - 1.5 * usq))).astype(np.float32)

    r2, u2, f2 = OD.stabilize(rho, u, solid, fd, feq)
    mon.emergency_stabilization(ctx)
    torch.cuda.synchronize()
    assert np.array_equal(_xyz(ctx.rho), r2, equal_nan=True)
    assert np.allclose(_uxyz(ctx.u), u2, rtol=1e-6, atol=1e-7)
    fd2 = np.ascontiguousarray(ctx.dense_populations().cpu().numpy().transpose(0, 3, 2, 1))
    fluid = solid == 0
    assert np.isfinite(fd2[:, fluid]).all()
    assert np.allclose(fd2[:, fluid], f2[:, fluid], rtol=1e-5, atol=1e-6)
    # the repaired state steps on and the monitor reads it clean
    ctx.step(20)
    ctx.macroscopic(macro=True)
    assert D.NumericalStabilityMonitor().check_field_stability(ctx) == 0


def test_lbm_diagnostics_updates_match_oracle():
    print("case: LBMDiagnostics three updates, V60 48^3 f32", flush=True)
    wl, ctx, solid = _state("f32", steps=50)
    diag = D.LBMDiagnostics(ctx)
    sc = diag.scales
    st = {"initial_mass": 0.0, "accumulated_inflow": 0.0, "prev_momentum": np.zeros(3)}
    prev = None
    for k in range(3):
        d = diag.update_diagnostics(step_num=100 * k, simulation_time=0.1 * k, force_update=True)
        rho, u = _xyz(ctx.rho), _uxyz(ctx.u)
        q = OD.lbm_quality(rho, u)
        for key, val in q.items():
            got = d["lbm_quality"][key]
            assert got == val if isinstance(val, str) else _close(got, val, rt=1e-9, at=1e-12), (key, got, val)
        c = OD.conservation(rho, u, st, gravity_lu=sc.gravity_lu, dt=sc.dt)
        for key in ("total_mass", "initial_mass", "mass_conservation_error", "momentum_z_change",
                    "expected_gravity_change", "conservation_grade"):
            got, val = d["conservation"][key], c[key]
            assert got == val if isinstance(val, str) else _close(got, val, rt=1e-9, at=1e-9), (key, got, val)
        t = OD.temporal_stability(rho, u, None if prev is None else prev[0], None if prev is None else prev[1])
        if prev is None:
            assert d["temporal_stability"] == t
        else:
            for key, val in t.items():
                got = d["temporal_stability"][key]
                assert got == val if isinstance(val, str) else _close(got, val, rt=1e-8, at=1e-12), (key, got, val)
        p = OD.v60_physics(rho, u, sc.scale_length, sc.scale_velocity)
        for key, val in p.items():
            assert _close(d["v60_physics"][key], val, rt=1e-9, at=1e-15), (key, d["v60_physics"][key], val)
        a = OD.flow_analysis(u)
        for key, val in a.items():
            assert _close(d["flow_analysis"][key], val, rt=1e-8, at=1e-12), (key, d["flow_analysis"][key], val)
        prev = (rho.astype(np.float64), u.astype(np.float64))
        ctx.step(40)
    rep = diag.get_summary_report()
    assert rep["history_length"] == 3 and rep["current_quality_grades"]["lbm_quality"] in ("Excellent", "Good", "Caution")


def test_export_matches_reference_bytes(tmp_path):
    print("case: analysis export 32^3 f32", flush=True)
    wl, ctx, solid = _state("f32", n=32, steps=30)
    files = D.export_data_for_analysis(ctx, simulation_time=1.5, step_num=30, out_dir=str(tmp_path))
    rho, u = _xyz(ctx.rho), _uxyz(ctx.u)
    sc = D.DiagnosticScales.reference(ctx.config.nz)
    meta = {"simulation_time": 1.5, "step_number": 30, "grid_size": [32, 32, 32],
            "physical_parameters": {"scale_length": sc.scale_length, "scale_time": sc.scale_time,
                                    "reynolds_number": 0.0}}
    want = OD.export_json_text(OD.export_data(u, rho, meta))
    got = open(files[0]).read()
    assert got == want
    z = np.load(files[1])
    assert sorted(z.files) == ["density_field", "pressure_field"]
    assert np.array_equal(z["pressure_field"], np.array((rho * (1.0 / 3.0)).tolist()))


def test_full_size_statistics():
    print("case: field statistics V60 224^3 f32", flush=True)
    wl, ctx, solid = _state("f32", n=224, steps=20)
    s, zp = D.field_statistics(ctx, populations=True)
    assert s["cells"] == 224 ** 3 and s["fluid"] == int((solid == 0).sum())
    assert abs(s["sum_drho"]) < 1e-3 * s["cells"] and s["fluid_pop_nan"] == 0 and s["fluid_pop_inf"] == 0
    assert abs(zp.sum() - (s["cells"] + s["sum_drho"])) < 1e-6 * s["cells"]
    json.dumps(D.visualizer_statistics(ctx))


def _box_ctx(nx, ny, nz, precision="f32"):
    """A closed box with a few solid cells inside (kind array [z, y, x])."""
    from pourover_amd.engine import SolverConfig
    from pourover_amd.geometry import link_flags
    kind = np.zeros((nz, ny, nx), np.uint8)
    kind[nz // 2, ny // 2, :] = L.POL_KIND_SOLID
    kind[0, 0, 0] = L.POL_KIND_SOLID
    cfg = SolverConfig(nx, ny, nz, precision=precision, tau=0.8, periodic=(False, False, False))
    ctx = HipContext(cfg, link_flags(kind)[0])
    solid = (kind == L.POL_KIND_SOLID).transpose(2, 1, 0).astype(np.uint8)
    return ctx, solid


def _expected(rho, u, solid, prev_rho=None, prev_u=None, phase=None):
    """The statistics of finite fields ([x, y, z] arrays) restated in numpy."""
    n = rho.size
    nz = rho.shape[2]
    r64, m = rho.astype(np.float64), OD.umag(u).astype(np.float64)
    mach = OD.umag(u) / OD.CS_F32
    e = {"cells": n, "sum_drho": np.sum(r64 - 1.0), "sum_drho2": np.sum((r64 - 1.0) ** 2),
         "max_abs_drho": np.max(np.abs(r64 - 1.0)), "sum_umag": m.sum(), "sum_umag2": (m * m).sum(),
         "max_umag": m.max(), "active": np.sum(mach > np.float32(1e-10)),
         "mach_gt": np.sum(mach > np.float32(0.1)),
         "sum_mach_active": mach[mach > np.float32(1e-10)].astype(np.float64).sum(),
         "umag_active": np.sum(m > np.float32(1e-10)), "sum_umag_active": m[m > np.float32(1e-10)].sum(),
         "outlet_down": np.maximum(-u[:, :, 0:5, 2].astype(np.float64), 0).sum()}
    mom = np.sum(r64[..., None] * u.astype(np.float64), axis=(0, 1, 2))
    e.update(mom_x=mom[0], mom_y=mom[1], mom_z=mom[2])
    for i, k in enumerate("xyz"):
        e["sum_u" + k] = u[..., i].astype(np.float64).sum()
        e["sum_u%s2" % k] = (u[..., i].astype(np.float64) ** 2).sum()
    zt, zm = int(nz * 0.8), int(nz * 0.3)
    e.update(sum_umag_top=m[:, :, zt:].sum(), sum_umag_mid=m[:, :, zm:zt].sum(), sum_umag_bot=m[:, :, :zm].sum())
    e["drho_prev2"] = 0.0 if prev_rho is None else np.sum((r64 - prev_rho.astype(np.float64)) ** 2)
    e["du_prev2"] = 0.0 if prev_u is None else np.sum((u.astype(np.float64) - prev_u.astype(np.float64)) ** 2)
    scan = OD.stability_scan(rho, u, solid)
    e.update(fluid=int(np.sum(solid == 0)), fluid_rho_min=scan["min_density"], fluid_rho_max=scan["max_density"],
             fluid_umax=scan["max_velocity"])
    vis = OD.visualizer_statistics(rho, u, solid, np.ones_like(rho) if phase is None else phase)
    e.update(int_cells=vis["total_nodes"], int_water=vis["total_water_mass"], int_air=vis["total_air_mass"],
             int_umax=vis["max_velocity"], int_umin=vis["min_velocity"],
             int_usum=vis["avg_velocity"] * vis["total_nodes"])
    return e, r64.sum(axis=(0, 1))


# (nx, ny, nz): 4-cell groups that never cross a row, groups that cross rows
# (nx = 6, a plane of 60 cells), one cell per lane (a plane of 323 cells; nx = 3)
@pytest.mark.parametrize("shape", [(8, 12, 10), (6, 10, 9), (17, 19, 11), (3, 8, 7)])
@pytest.mark.parametrize("extra", [False, True])
def test_statistics_kernel_forms(shape, extra):
    """Every statistic of random finite fields on boxes that take each form of
    the statistics kernel (four cells per lane with and without row-crossing
    groups, one cell per lane), with and without a previous state and phase,
    against the numpy restatement."""
    print("case: statistics kernel forms %s extra=%s" % (shape, extra), flush=True)
    ctx, solid = _box_ctx(*shape)
    rng = np.random.default_rng(sum(shape) + extra)
    nx, ny, nz = shape
    r = rng.uniform(0.9, 1.1, (nz, ny, nx)).astype(np.float32)
    uu = rng.normal(0, 0.05, (3, nz, ny, nx)).astype(np.float32)
    uu[:, 1, 1, 1] = 0.0                     # |u| = +0: inactive
    uu[:, 2, 2, 1] = np.float32(1e-12)       # under both activity thresholds
    ctx.rho.copy_(torch.from_numpy(r))
    ctx.u.copy_(torch.from_numpy(uu))
    prev_rho = prev_u = phase = None
    if extra:
        prev_rho = ctx.rho * 0.999 + 0.001
        prev_u = ctx.u * 0.9
        phase = torch.from_numpy(rng.uniform(-0.5, 1.0, (nz, ny, nx)).astype(np.float32)).cuda()
    s, zp = D.field_statistics(ctx, prev_rho, prev_u, phase)
    rho, u = _xyz(ctx.rho), _uxyz(ctx.u)
    e, ze = _expected(rho, u, solid, None if prev_rho is None else _xyz(prev_rho),
                      None if prev_u is None else _uxyz(prev_u), None if phase is None else _xyz(phase))
    exact = {"cells", "max_abs_drho", "max_umag", "active", "mach_gt", "umag_active", "fluid", "fluid_rho_min",
             "fluid_rho_max", "fluid_umax", "int_cells", "int_umax", "int_umin"}
    for k, v in e.items():
        if k in exact:
            assert s[k] == v, (k, s[k], v)
        else:
            assert _close(s[k], v, at=1e-12 * rho.size), (k, s[k], v)
    assert np.allclose(zp, ze, rtol=RT, atol=0)
    assert (s["fluid_rho_nan"], s["fluid_rho_inf"], s["fluid_u_nan"], s["fluid_u_inf"]) == (0, 0, 0, 0)


def test_statistics_nonfinite_velocity():
    """|u| extremes and counts with NaN and infinite velocities: max |u| is
    NaN when any |u| is; the non-solid maximum keeps finite values only; the
    interior extremes skip NaN (they keep +inf) and give their identities
    when every interior value is NaN."""
    print("case: statistics with NaN / inf velocities", flush=True)
    nx, ny, nz = 8, 8, 8
    ctx, solid = _box_ctx(nx, ny, nz)
    base = np.full((3, nz, ny, nx), 0.01, np.float32)
    base[0, 3, 3, 3] = 0.2                   # the largest finite |u|, interior
    base[1, 0, 4, 4] = 0.3                   # larger, on the z = 0 face (not interior)
    ctx.rho.fill_(1.0)
    ctx.u.copy_(torch.from_numpy(base))
    u = base.copy()
    u[2, 2, 5, 5] = np.nan                   # interior, non-solid
    u[0, 6, 2, 2] = -np.inf                  # interior, non-solid: |u| = +inf
    u[1, 7, 7, 7] = np.nan                   # boundary face
    ctx.u.copy_(torch.from_numpy(u))
    s, _ = D.field_statistics(ctx)
    assert np.isnan(s["max_umag"]) and np.isnan(s["sum_umag"])
    assert (s["fluid_u_nan"], s["fluid_u_inf"]) == (2, 1)
    assert s["fluid_umax"] == float(OD.umag(base[:, 0, 4, 4][None, :])[0])
    assert s["int_umax"] == np.inf
    assert s["int_umin"] == float(OD.umag(np.full((1, 3), 0.01, np.float32))[0])
    # every interior cell NaN: the interior extremes keep their identities
    u2 = base.copy()
    u2[0, 1:nz - 1, 1:ny - 1, 1:nx - 1] = np.nan
    ctx.u.copy_(torch.from_numpy(u2))
    s, _ = D.field_statistics(ctx)
    assert s["int_umax"] == -1e300 and s["int_umin"] == 1e300
    assert s["int_cells"] == int(np.sum(solid[1:nx - 1, 1:ny - 1, 1:nz - 1] == 0))
    assert s["fluid_umax"] == float(OD.umag(base[:, 0, 4, 4][None, :])[0])
