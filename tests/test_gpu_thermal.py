"""Thermal D3Q7 on the device (csrc/thermal.hip through pol_thermal_*, and
pourover_amd.thermal) against the numpy restatement of the reference's
thermal solver (oracle/thermal.py) on the same seeded fields.

The kernels evaluate the reference's expressions operation by operation with
contraction off, so populations, temperatures, heat sources and fluxes
compare bit for bit (FP32 and FP64).  The mean temperature is an FP64 sum in
block order and compares to rtol 1e-12."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import thermal as OT
from pourover_amd import thermal as TH
from pourover_amd.engine import SolverConfig
from pourover_amd.solver import HipLBMSolver

pytestmark = pytest.mark.gpu

NX, NY, NZ = 24, 20, 16


def _np_dtype(precision):
    return np.float64 if precision == "f64" else np.float32


def _fields(seed, dtype, hot=(20.0, 90.0)):
    rng = np.random.default_rng(seed)
    T = rng.uniform(*hot, size=(NX, NY, NZ)).astype(dtype)
    hs = rng.uniform(-0.05, 0.05, size=(NX, NY, NZ)).astype(dtype)
    u = rng.uniform(-0.05, 0.05, size=(NX, NY, NZ, 3)).astype(dtype)
    return T, hs, u


def _solver(precision, omega=None):
    th = TH.ThermalLBM(grid=(NX, NY, NZ), precision=precision)
    if omega is not None:
        th.omega_thermal = omega
        th.tau_thermal = 1.0 / omega
        th.enable_convection_coupling(th.enable_convection)  # pushes the new omega
    return th


def _g_host(th):
    return th.g  # [nx, ny, nz, 7]


def _assert_same(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    assert a.dtype == b.dtype, (what, a.dtype, b.dtype)
    it = np.int64 if a.dtype == np.float64 else np.int32
    bad = np.argwhere(a.view(it) != b.view(it))  # bit patterns
    if bad.size:
        i = tuple(bad[0])
        raise AssertionError("%s differs at %d places, first %s: %r vs %r" % (what, len(bad), i, a[i], b[i]))


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("convection", [False, True])
def test_thermal_steps_match_oracle(precision, convection):
    print("case: thermal %dx%dx%d %s convection=%s" % (NX, NY, NZ, precision, convection), flush=True)
    dt = _np_dtype(precision)
    T, hs, u = _fields(11 + convection, dt)
    th = _solver(precision, omega=1.3)
    th.initialize_from(T)
    th.heat_source.from_numpy(hs)
    if convection:
        th.enable_convection_coupling(True)
        th._vel.copy_(torch.from_numpy(np.ascontiguousarray(u.transpose(3, 2, 1, 0))))
    st = {"g": OT.equilibrium(T).astype(dt), "T": T, "heat_source": hs, "u": u}
    _assert_same(_g_host(th), st["g"], "initial g")
    steps = 25
    for n in range(steps):
        st = OT.step(st, 1.3, convection=convection)
        if n == 0:
            assert th.step()
        elif n == 1:
            assert th.run(steps - 1)
    torch.cuda.synchronize()
    _assert_same(_g_host(th), st["g"], "g after %d steps" % steps)
    _assert_same(th.temperature.to_numpy(), st["T"], "T")
    _assert_same(th.temperature_old.to_numpy(), st["T_old"], "T_old")
    _assert_same(th.heat_source.to_numpy(), st["heat_source"], "heat source")
    assert th.current_step == steps


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_heat_flux_and_stats(precision):
    print("case: thermal flux/stats %s" % precision, flush=True)
    dt = _np_dtype(precision)
    T, hs, _ = _fields(3, dt)
    th = _solver(precision, omega=0.9)
    th.initialize_from(T)
    th.heat_source.from_numpy(hs)
    th.run(3)
    st = {"g": OT.equilibrium(T).astype(dt), "T": T, "heat_source": hs}
    for _ in range(3):
        st = OT.step(st, 0.9)
    k = np.full((NX, NY, NZ), 0.68, dt)
    flux = th.heat_flux.to_numpy()
    _assert_same(flux, OT.heat_flux(st["T"], k), "heat flux")
    mn, mx, mean = th.get_temperature_stats()
    assert mn == float(st["T"].min()) and mx == float(st["T"].max())
    ref_mean = float(np.sum(st["T"].astype(np.float64)) / st["T"].size)
    assert abs(mean - ref_mean) <= 1e-12 * abs(ref_mean)


@pytest.mark.parametrize("kind", [1, 2, 3])
def test_boundary_forms(kind):
    print("case: thermal boundary kind %d" % kind, flush=True)
    T, hs, _ = _fields(7, np.float32)
    th = _solver("f32", omega=1.1)
    th.initialize_from(T)
    th.heat_source.from_numpy(hs)
    th.run(2)
    st = {"g": OT.equilibrium(T), "T": T, "heat_source": hs}
    for _ in range(2):
        st = OT.step(st, 1.1)
    rng = np.random.default_rng(kind)
    mask = rng.random((NX, NY, NZ)) < 0.2
    mask[:, :, 0] = True
    if kind == 1:
        th.apply_dirichlet_bc(mask, 93.5)
        st = OT.boundary(st, 1, mask, 93.5)
    elif kind == 2:
        th.apply_neumann_bc(mask, -0.25)
        st = OT.boundary(st, 2, mask, -0.25)
    else:
        th.apply_convective_bc(mask, 0.3, 25.0)
        st = OT.boundary(st, 3, mask, 0.3, 25.0)
    _assert_same(_g_host(th), st["g"], "g")
    _assert_same(th.temperature.to_numpy(), st["T"], "T")
    _assert_same(th.heat_source.to_numpy(), st["heat_source"], "heat source")
    # and one step further from the changed state
    th.step()
    st = OT.step(st, 1.1)
    _assert_same(_g_host(th), st["g"], "g after a step")


def test_base_source_reset_without_convection():
    """POL_THERMAL_RESET_SOURCE alone: every step uses S = base and writes it."""
    print("case: thermal base source reset", flush=True)
    T, hs, _ = _fields(13, np.float32)
    th = _solver("f32", omega=1.2)
    th.initialize_from(T)
    th.heat_source.from_numpy(hs * 7)
    th.use_base_heat_source(hs)
    th.run(4)
    st = {"g": OT.equilibrium(T), "T": T, "heat_source": hs}
    for _ in range(4):
        st = OT.step(st, 1.2)
    _assert_same(_g_host(th), st["g"], "g")
    _assert_same(th.heat_source.to_numpy(), hs, "heat source = base")
    th.use_base_heat_source(None, enable=False)
    th.heat_source.from_numpy(hs * 2)
    th.step()
    st = OT.step(dict(st, heat_source=hs * 2), 1.2)
    _assert_same(_g_host(th), st["g"], "g after the base is released")


def test_unstable_cells_counted():
    print("case: thermal stability test", flush=True)
    T, hs, _ = _fields(5, np.float32)
    T[3, 4, 5] = 1000.0  # above the 150 limit, still above it after one step
    T[10, 10, 10] = np.nan
    th = _solver("f32", omega=1.0)
    th.initialize_from(T)
    th.heat_source.from_numpy(hs)
    st = OT.step({"g": OT.equilibrium(T), "T": T, "heat_source": hs}, 1.0)
    assert th.step() is False
    assert th._last_unstable == OT.unstable_cells(st["T"], st["g"]) > 0
    assert th.check_numerical_stability() == 1
    assert th.current_step == 0


def test_reference_constructor_and_initialisation():
    print("case: thermal constructor / init_temperature_field", flush=True)
    th = TH.ThermalLBM(grid=(NX, NY, NZ))
    alpha_lu = 1.6e-7 * 0.0625 / 0.000625 ** 2
    assert th.tau_thermal == pytest.approx(alpha_lu * 3.0 + 0.5, rel=1e-15)
    assert th.get_effective_thermal_diffusivity() == pytest.approx(1.6e-7, rel=1e-12)
    assert not th.step()  # not initialised
    th.complete_initialization(25.0, 93.0, 5)
    T = th.temperature.to_numpy()
    assert np.all(T[:, :, :5] == np.float32(93.0)) and np.all(T[:, :, 5:] == np.float32(25.0))
    _assert_same(_g_host(th), OT.equilibrium(T), "g = w T")
    assert th.step() and th.current_step == 1
    clamp = TH.ThermalLBM(thermal_diffusivity=1e-12, grid=(NX, NY, NZ))
    assert clamp.tau_thermal == 0.51


@pytest.mark.parametrize("subcycles", [1, 2])
def test_coupled_step_matches_oracle(subcycles):
    """thermal_fluid_coupled.py:143-209: fluid step, velocity hand-over,
    heat-source reset to its base, thermal sub-cycles -- two coupled steps.
    One sub-cycle runs the reset inside the thermal step; two run it as a
    copy, the source accumulating over the sub-cycles."""
    print("case: thermal-fluid coupled step, %d sub-cycles" % subcycles, flush=True)
    cfg = SolverConfig(NX, NY, NZ, precision="f32", tau=0.8, gravity=(2e-5, 0.0, -1e-5))
    fluid = HipLBMSolver(cfg)
    fluid.initialize_fields()
    th = TH.ThermalLBM(solver=fluid)
    T, hs, _ = _fields(9, np.float32)
    th.initialize_from(T)
    base = hs.copy()
    cs = TH.ThermalFluidCoupledSolver(TH.CouplingConfig(thermal_subcycles=subcycles), fluid_solver=fluid,
                                      thermal_solver=th)
    cs.set_base_heat_source(base)
    cs.is_initialized = True
    assert cs.fused_reset == (subcycles == 1)
    st = {"g": OT.equilibrium(T), "T": T, "heat_source": base}
    for _ in range(2):
        assert cs.step()
        u = fluid.u.to_numpy()
        st = dict(st, heat_source=base.copy(), u=u)
        for _ in range(subcycles):
            st = OT.step(st, th.omega_thermal, convection=True)
    _assert_same(_g_host(th), st["g"], "coupled g")
    _assert_same(th.temperature.to_numpy(), st["T"], "coupled T")
    _assert_same(th.heat_source.to_numpy(), st["heat_source"], "coupled heat source")
    assert np.abs(u).max() > 0


def test_coupled_solver_reference_interface():
    """thermal_fluid_coupled.py:58-355: the solver builds its own fluid and
    thermal solvers, initialize_system, steps, diagnostics, state, reset."""
    print("case: coupled solver interface", flush=True)
    cs = TH.ThermalFluidCoupledSolver(grid=(NX, NY, NZ))
    assert not cs.step()                                   # not initialised
    assert cs.get_coupling_diagnostics() == {"status": "not_started"}
    base = np.full((NX, NY, NZ), 0.01, np.float32)
    cs.initialize_system({}, {"T_initial": 25.0, "T_hot_region": 90.0, "hot_region_height": 4}, base)
    for _ in range(3):
        assert cs.step()
    d = cs.get_coupling_diagnostics()
    # over-relaxation (omega = 1.73) undershoots next to the 90 / 25 step; the
    # coupled stability bounds are -10 and 150
    assert d["coupling_step"] == 3 and -10.0 < d["thermal_stats"]["T_min"] < d["thermal_stats"]["T_max"] < 150.0
    assert d["coupling_config"]["convection_enabled"]
    st = cs.save_coupling_state(3)
    assert st["temperature"].shape == (NX, NY, NZ) and st["velocity"].shape == (NX, NY, NZ, 3)
    _assert_same(cs.thermal_solver.heat_source.to_numpy(), base, "heat source = base after the fused reset")
    cs.reset_coupling_system()
    assert cs.coupling_step == 0 and not cs.is_initialized
