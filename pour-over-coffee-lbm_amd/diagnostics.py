"""Field diagnostics of the HIP solver: stability monitor, LBM diagnostics,
visualiser statistics and the analysis export.

Mirrors (reference):
  src/core/numerical_stability.py:18-299   NumericalStabilityMonitor,
                                           ErrorRecoverySystem
  src/visualization/lbm_diagnostics.py:13-521  CircularBuffer, LBMDiagnostics
  src/visualization/visualizer.py:120-193  interior statistics
  src/visualization/enhanced_visualizer.py:1487-1559  export_data_for_analysis
with the same method names, thresholds, dictionary keys and grades.  The
reference pulls rho, u (and f) to the host and reduces them in numpy or Taichi
loops; here every number comes from one device pass (pol_field_stats, FP64,
deterministic) and only the ~40 statistics cross to the host.  Solvers are
HipLBMSolver instances (or anything with a HipContext `ctx`).

Units follow the reference: lattice units, with the physical scales of
config/core.py:27-81 and config/physics.py:175-177 (DiagnosticScales) where
the reference multiplies by them.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import time
from collections import deque
from dataclasses import dataclass
from datetime import datetime
from typing import Optional

import numpy as np
import torch

from . import _lib as L


@dataclass
class DiagnosticScales:
    """Physical scales of the reference configuration (config/core.py:27-81,
    config/physics.py:33,175-177) for a grid of height nz."""

    scale_length: float
    scale_velocity: float
    scale_time: float
    dt: float
    cs2: float
    gravity_lu: float

    @staticmethod
    def reference(nz: int = 224) -> "DiagnosticScales":
        length = 0.14 / nz
        vel = 0.01
        t = length / vel * 1.2
        return DiagnosticScales(length, vel, t, 1.0, 1.0 / 3.0, 9.81 * t * t / length * 0.5)


def _ctx(solver):
    return solver.ctx if hasattr(solver, "ctx") else solver


def _ensure_macro(solver) -> None:
    """A HipLBMSolver refreshes rho / u when the populations moved on since
    they were last computed; a bare HipContext is read as it stands (call
    ctx.macroscopic first)."""
    if hasattr(solver, "_ensure_macro"):
        solver._ensure_macro()


def field_statistics(solver, prev_rho: Optional[torch.Tensor] = None, prev_u: Optional[torch.Tensor] = None,
                     phase: Optional[torch.Tensor] = None, populations: bool = False):
    """(dict of the POL_STAT_* slots, z profile sum_{x,y} rho per plane)."""
    _ensure_macro(solver)
    ctx = _ctx(solver)
    cfg = ctx.config
    o = L.PolStatsOpts()
    o.prev_rho = 0 if prev_rho is None else prev_rho.data_ptr()
    o.prev_u = 0 if prev_u is None else prev_u.data_ptr()
    o.phase = 0 if phase is None else phase.data_ptr()
    o.z_top, o.z_mid = int(cfg.nz * 0.8), int(cfg.nz * 0.3)
    o.z_outlet = 5
    o.with_populations = 1 if populations else 0
    out = np.zeros(L.POL_NSTAT, dtype=np.float64)
    zp = np.zeros(cfg.nz, dtype=np.float64)
    st = ctx.lib.pol_field_stats(ctx.handle, C.byref(o), out.ctypes.data, zp.ctypes.data, L.stream_handle())
    L.check(st, ctx.handle, "pol_field_stats")
    return dict(zip(L.STAT_SLOTS, out.tolist())), zp


def _std(sum_x: float, sum_x2: float, n: float) -> float:
    m = sum_x / n
    return math.sqrt(max(sum_x2 / n - m * m, 0.0))


# ---- numerical_stability.py -------------------------------------------------------------
class NumericalStabilityMonitor:
    """numerical_stability.py:18-243: scan of the non-solid cells for NaN / Inf
    density, velocity and populations, the density range and max |u|."""

    def __init__(self):
        self.MAX_VELOCITY_THRESHOLD = 0.3
        self.MIN_DENSITY_THRESHOLD = 0.1
        self.MAX_DENSITY_THRESHOLD = 5.0
        self.MAX_NAN_TOLERANCE = 10
        self.error_history: list = []
        self.consecutive_errors = 0
        self._last: Optional[dict] = None

    def check_field_stability(self, solver) -> int:
        """0 normal, 1 warning, 2 severe (numerical_stability.py:53-110)."""
        s, _ = field_statistics(solver, phase=None, populations=True)
        nan = int(s["fluid_rho_nan"] + s["fluid_u_nan"] + s["fluid_pop_nan"])
        inf = int(s["fluid_rho_inf"] + s["fluid_u_inf"] + s["fluid_pop_inf"])
        finite_rho = s["fluid"] - s["fluid_rho_nan"] - s["fluid_rho_inf"] > 0
        self._last = {"max_velocity": float(s["fluid_umax"]),
                      "min_density": float(s["fluid_rho_min"]) if finite_rho else 1e10,
                      "max_density": float(s["fluid_rho_max"]) if finite_rho else 0.0,
                      "nan_count": nan, "inf_count": inf}
        t = self._last
        if (nan > 0 or inf > 0 or t["max_velocity"] > self.MAX_VELOCITY_THRESHOLD
                or t["min_density"] < self.MIN_DENSITY_THRESHOLD or t["max_density"] > self.MAX_DENSITY_THRESHOLD):
            if nan > self.MAX_NAN_TOLERANCE or t["max_velocity"] > 0.5:
                return 2
            return 1
        return 0

    def check_stability(self, solver) -> bool:
        try:
            return self.check_field_stability(solver) == 0
        except Exception:
            return False

    def diagnose_stability(self, solver, step: int) -> dict:
        try:
            status = self.check_field_stability(solver)
            t = self._last
            report = {"step": step, "status": status, "max_velocity": t["max_velocity"],
                      "min_density": t["min_density"], "max_density": t["max_density"],
                      "nan_count": t["nan_count"], "inf_count": t["inf_count"],
                      "mach_number": t["max_velocity"] / np.sqrt(1.0 / 3.0), "is_stable": status == 0}
            if status > 0:
                self.consecutive_errors += 1
                self.error_history.append((step, status))
                if len(self.error_history) > 100:
                    self.error_history.pop(0)
            else:
                self.consecutive_errors = 0
            return report
        except Exception as e:  # the reference reports a failed check as severe
            return {"step": step, "status": 2, "error": str(e), "is_stable": False}

    def emergency_stabilization(self, solver) -> None:
        """numerical_stability.py:165-197 on the device (pol_stabilize)."""
        _ensure_macro(solver)
        ctx = _ctx(solver)
        st = ctx.lib.pol_stabilize(ctx.handle, self.MAX_VELOCITY_THRESHOLD, self.MIN_DENSITY_THRESHOLD,
                                   self.MAX_DENSITY_THRESHOLD, L.stream_handle())
        L.check(st, ctx.handle, "pol_stabilize")

    def get_statistics(self) -> dict:
        if self._last is None:
            return {"max_velocity": 0.0, "min_density": 1.0, "max_density": 1.0, "nan_count": 0, "inf_count": 0}
        return dict(self._last)

    def print_stability_report(self, report: dict) -> None:
        step, status = report["step"], report["status"]
        if status == 0:
            print(f"Step {step}: stable")
        elif status == 1:
            print(f"Step {step}: warning")
            print(f"   Max velocity: {report['max_velocity']:.6f}")
            print(f"   Density range: [{report['min_density']:.3f}, {report['max_density']:.3f}]")
            if report["nan_count"] > 0:
                print(f"   NaN count: {report['nan_count']}")
        else:
            print(f"Step {step}: diverging")
            print(f"   Max velocity: {report['max_velocity']:.6f}")
            print(f"   NaN/Inf: {report['nan_count']}/{report['inf_count']}")
            print(f"   consecutive errors: {self.consecutive_errors}")

    def should_abort_simulation(self) -> bool:
        return self.consecutive_errors > 10


class ErrorRecoverySystem:
    """numerical_stability.py:246-299."""

    def __init__(self):
        self.monitor = NumericalStabilityMonitor()
        self.recovery_attempts = 0
        self.max_recovery_attempts = 3

    def step_with_monitoring(self, solver, step: int) -> dict:
        try:
            solver.step()
            report = self.monitor.diagnose_stability(solver, step)
            if report["status"] == 2:
                self._attempt_recovery(solver, step)
                self.recovery_attempts += 1
            elif report["status"] == 1:
                if step % 100 == 0:
                    self.monitor.print_stability_report(report)
            else:
                self.recovery_attempts = 0
            if self.monitor.should_abort_simulation():
                raise RuntimeError(f"simulation aborted at step {step}: consecutive divergence")
            return report
        except Exception:
            if self.recovery_attempts < self.max_recovery_attempts:
                self.recovery_attempts += 1
                self._attempt_recovery(solver, step)
                return self.step_with_monitoring(solver, step)
            raise

    def _attempt_recovery(self, solver, step: int) -> None:
        self.monitor.emergency_stabilization(solver)
        if hasattr(solver, "swap_fields"):
            solver.swap_fields()  # the fused step ping-pongs its buffers itself: a no-op here


# ---- lbm_diagnostics.py -------------------------------------------------------------
class CircularBuffer:
    """lbm_diagnostics.py:13-31."""

    def __init__(self, max_size: int = 1000):
        self.data: deque = deque(maxlen=max_size)
        self.timestamps: deque = deque(maxlen=max_size)

    def add(self, timestamp, data_dict) -> None:
        self.timestamps.append(timestamp)
        self.data.append(data_dict)

    def get_recent(self, n: int = 10):
        return list(self.timestamps)[-n:], list(self.data)[-n:]

    def get_all(self):
        return list(self.timestamps), list(self.data)


class LBMDiagnostics:
    """lbm_diagnostics.py:33-521 on device statistics.  The previous step's
    rho / u stay on the device (one copy each) for the temporal-stability
    norms instead of host copies."""

    def __init__(self, lbm_solver, multiphase=None, particles=None, pouring=None, filter_system=None,
                 scales: Optional[DiagnosticScales] = None):
        self.lbm = lbm_solver
        self.multiphase = multiphase
        self.particles = particles
        self.pouring = pouring
        self.filter = filter_system
        cfg = _ctx(lbm_solver).config
        self.scales = scales or DiagnosticScales.reference(cfg.nz)
        self.history = CircularBuffer(max_size=1000)
        self.detailed_history = CircularBuffer(max_size=100)
        self.light_monitoring_freq = 5
        self.medium_monitoring_freq = 10
        self.heavy_monitoring_freq = 100
        self.prev_rho: Optional[torch.Tensor] = None
        self.prev_u: Optional[torch.Tensor] = None
        self.prev_momentum = np.zeros(3)
        self.initial_mass = 0.0
        self.accumulated_inflow = 0.0
        self.reference_density = 1.0
        self.calculation_times = {"light": [], "medium": [], "heavy": []}
        self._stats: Optional[dict] = None
        self._zprof: Optional[np.ndarray] = None

    # one device pass per diagnostic update, shared by every analysis below
    def _refresh(self) -> None:
        self._stats, self._zprof = field_statistics(self.lbm, self.prev_rho, self.prev_u)

    def _s(self) -> dict:
        if self._stats is None:
            self._refresh()
        return self._stats

    def adaptive_monitoring_frequency(self, step_num: int) -> None:
        if step_num < 100:
            self.light_monitoring_freq, self.medium_monitoring_freq = 2, 5
        elif step_num < 500:
            self.light_monitoring_freq, self.medium_monitoring_freq = 5, 10
        else:
            self.light_monitoring_freq, self.medium_monitoring_freq = 10, 20

    def update_diagnostics(self, step_num: int, simulation_time: float, force_update: bool = False) -> dict:
        self.adaptive_monitoring_frequency(step_num)
        self._refresh()
        d: dict = {"step": step_num, "time": simulation_time, "timestamp": datetime.now().isoformat()}
        if step_num % self.light_monitoring_freq == 0 or force_update:
            t0 = time.time()
            d.update({"temporal_stability": self.calculate_temporal_stability(),
                      "lbm_quality": self.analyze_lbm_numerical_quality()})
            self.calculation_times["light"].append(time.time() - t0)
        if step_num % self.medium_monitoring_freq == 0 or force_update:
            t0 = time.time()
            d.update({"conservation": self.check_conservation_laws(), "multiphase": self.analyze_multiphase_quality()})
            self.calculation_times["medium"].append(time.time() - t0)
        if step_num % self.heavy_monitoring_freq == 0 or force_update:
            t0 = time.time()
            d.update({"v60_physics": self.track_v60_physics(), "flow_analysis": self.detailed_flow_analysis(),
                      "diagnostics_performance": self.get_performance_stats()})
            self.detailed_history.add(simulation_time, d)
            self.calculation_times["heavy"].append(time.time() - t0)
        self.history.add(simulation_time, d)
        self._update_previous_step_data()
        return d

    def calculate_temporal_stability(self) -> dict:
        s = self._s()
        if self.prev_rho is None or self.prev_u is None:
            return {"status": "first_step"}
        n = s["cells"]
        rho_norm = math.sqrt(n + 2.0 * s["sum_drho"] + s["sum_drho2"])
        u_norm = math.sqrt(s["sum_umag2"])
        dr, du = math.sqrt(s["drho_prev2"]), math.sqrt(s["du_prev2"])
        st = {"density_change_rate": dr / self.scales.dt, "relative_density_change": dr / rho_norm,
              "velocity_change_rate": du / self.scales.dt, "relative_velocity_change": du / (u_norm + 1e-10)}
        if st["relative_density_change"] < 1e-6 and st["relative_velocity_change"] < 1e-4:
            st["stability_grade"] = "Excellent"
        elif st["relative_density_change"] < 1e-4 and st["relative_velocity_change"] < 1e-2:
            st["stability_grade"] = "Good"
        else:
            st["stability_grade"] = "Monitoring"
        return st

    def analyze_lbm_numerical_quality(self) -> dict:
        s = self._s()
        n = s["cells"]
        cs = float(np.float32(1.0 / np.sqrt(3.0)))
        q = {"max_mach": float(np.float32(s["max_umag"]) / np.float32(cs)),
             "mean_mach": s["sum_mach_active"] / s["active"] if s["active"] > 0 else 0.0,
             "mach_violation_ratio": s["mach_gt"] / n,
             "max_density_deviation": s["max_abs_drho"],
             "mean_density": 1.0 + s["sum_drho"] / n}
        q["density_variation_coeff"] = _std(s["sum_drho"], s["sum_drho2"], n) / q["mean_density"]
        if q["max_mach"] < 0.1 and q["max_density_deviation"] < 0.1:
            q["lbm_grade"] = "Excellent"
        elif q["max_mach"] < 0.2 and q["max_density_deviation"] < 0.2:
            q["lbm_grade"] = "Good"
        else:
            q["lbm_grade"] = "Caution"
        return q

    def check_conservation_laws(self) -> dict:
        s = self._s()
        mass = s["cells"] + s["sum_drho"]
        if self.initial_mass == 0.0:
            self.initial_mass = mass
        if self.pouring is not None and hasattr(self.pouring, "get_current_flow_rate"):
            try:
                self.accumulated_inflow += self.pouring.get_current_flow_rate() * self.scales.dt
            except Exception:
                pass
        err = abs(mass - (self.initial_mass + self.accumulated_inflow))
        c = {"total_mass": mass, "initial_mass": self.initial_mass, "accumulated_inflow": self.accumulated_inflow,
             "mass_conservation_error": err, "relative_mass_error": err / self.initial_mass}
        mom = np.array([s["mom_x"], s["mom_y"], s["mom_z"]])
        change = mom - self.prev_momentum
        g = self.scales.gravity_lu * mass * self.scales.dt
        c.update({"total_momentum": mom.tolist(), "momentum_z_change": float(change[2]),
                  "expected_gravity_change": float(g), "momentum_conservation_error": float(abs(change[2] - g))})
        c["conservation_grade"] = ("Excellent" if c["relative_mass_error"] < 1e-6 else
                                   "Good" if c["relative_mass_error"] < 1e-4 else "Moderate")
        self.prev_momentum = mom.copy()
        return c

    def analyze_multiphase_quality(self) -> dict:
        # the hot path is single-phase (the multiphase producer is not part of it)
        return {"status": "no_multiphase"}

    def track_v60_physics(self) -> dict:
        s = self._s()
        cfg = _ctx(self.lbm).config
        prof = self._zprof / (cfg.nx * cfg.ny)
        wet = np.where(prof > self.reference_density + 0.1)[0]
        L_ = self.scales.scale_length
        p = {"wetting_front_position": float(np.max(wet) * L_) if wet.size else 0.0,
             "system_water_volume": float(max(0.0, s["sum_drho"] * L_ ** 3))}
        inlet = 0.0
        if self.pouring is not None and hasattr(self.pouring, "get_current_flow_rate"):
            try:
                inlet = self.pouring.get_current_flow_rate()
            except Exception:
                pass
        p["inlet_flow_rate"] = float(inlet)
        p["outlet_flow_rate"] = self._estimate_outlet_flow()
        if self.particles is not None:
            p["particle_dynamics"] = self._analyze_coffee_bed_dynamics()
        return p

    def detailed_flow_analysis(self) -> dict:
        s = self._s()
        cfg = _ctx(self.lbm).config
        n = s["cells"]
        a = {"max_velocity": s["max_umag"],
             "mean_velocity": s["sum_umag_active"] / s["umag_active"] if s["umag_active"] > 0 else 0.0,
             "velocity_std": _std(s["sum_umag"], s["sum_umag2"], n),
             "flow_complexity": float(np.linalg.norm([_std(s["sum_u" + c], s["sum_u" + c + "2"], n) for c in "xyz"]))}
        plane = cfg.nx * cfg.ny
        zt, zm = int(cfg.nz * 0.8), int(cfg.nz * 0.3)
        a["top_region_flow"] = s["sum_umag_top"] / (plane * (cfg.nz - zt))
        a["middle_region_flow"] = s["sum_umag_mid"] / (plane * (zt - zm))
        a["bottom_region_flow"] = s["sum_umag_bot"] / (plane * zm) if zm > 0 else float("nan")
        return a

    def _analyze_coffee_bed_dynamics(self) -> dict:
        ps = self.particles
        pos = getattr(ps, "pos", None)
        act = getattr(ps, "active", None)
        if pos is None or act is None:
            return {"status": "no_particle_data"}
        p = pos.double().cpu().numpy().T  # [P, 3]
        a = act.cpu().numpy()
        p = p[a == 1]
        if len(p) == 0:
            return {"status": "no_active_particles"}
        cfg = _ctx(self.lbm).config
        L_ = self.scales.scale_length
        z = p[:, 2]
        r = np.sqrt((p[:, 0] - cfg.nx * 0.5) ** 2 + (p[:, 1] - cfg.ny * 0.5) ** 2)
        return {"active_particle_count": len(p), "bed_height_mean": float(np.mean(z) * L_),
                "bed_height_std": float(np.std(z) * L_),
                "bed_height_range": [float(np.min(z) * L_), float(np.max(z) * L_)],
                "radial_distribution_mean": float(np.mean(r) * L_), "radial_distribution_std": float(np.std(r) * L_)}

    def _estimate_outlet_flow(self) -> float:
        s = self._s()
        return float(s["outlet_down"] * self.scales.scale_velocity * self.scales.scale_length ** 2)

    def _update_previous_step_data(self) -> None:
        ctx = _ctx(self.lbm)
        if self.prev_rho is None:
            self.prev_rho = torch.empty_like(ctx.rho)
            self.prev_u = torch.empty_like(ctx.u)
        self.prev_rho.copy_(ctx.rho)
        self.prev_u.copy_(ctx.u)
        self._stats = None

    def get_performance_stats(self) -> dict:
        out = {}
        for k in ("light", "medium", "heavy"):
            t = self.calculation_times[k]
            out[f"{k}_calc_count"] = len(t)
            out[f"{k}_avg_time"] = float(np.mean(t)) if t else 0.0
            out[f"{k}_total_time"] = float(np.sum(t)) if t else 0.0
        return out

    def get_current_diagnostics(self) -> dict:
        return self.history.data[-1] if self.history.data else {}

    def get_diagnostics_history(self, detailed: bool = False):
        return self.detailed_history.get_all() if detailed else self.history.get_all()

    def get_summary_report(self) -> dict:
        cur = self.get_current_diagnostics()
        perf = self.get_performance_stats()
        return {"monitoring_status": "active",
                "total_calculations": sum(perf.get(f"{t}_calc_count", 0) for t in ("light", "medium", "heavy")),
                "total_diagnostic_time": sum(perf.get(f"{t}_total_time", 0) for t in ("light", "medium", "heavy")),
                "history_length": len(self.history.data),
                "current_quality_grades": {
                    "stability": cur.get("temporal_stability", {}).get("stability_grade", "Unknown"),
                    "lbm_quality": cur.get("lbm_quality", {}).get("lbm_grade", "Unknown"),
                    "conservation": cur.get("conservation", {}).get("conservation_grade", "Unknown")}}


# ---- visualizer.py:120-193 ---------------------------------------------------------------
def visualizer_statistics(solver) -> dict:
    """The interior (1 <= i <= n - 2) non-solid statistics of
    visualizer.py:120-167 (get_statistics keys)."""
    phase = getattr(solver, "phase", None)
    ph = phase.to_torch() if phase is not None and hasattr(phase, "to_torch") else None
    s, _ = field_statistics(solver, phase=ph)
    n = int(s["int_cells"])
    return {"total_water_mass": s["int_water"], "total_air_mass": s["int_air"],
            "max_velocity": s["int_umax"] if n else 0.0, "min_velocity": s["int_umin"] if n else 0.0,
            "avg_velocity": s["int_usum"] / n if n else 0.0, "total_nodes": n}


# ---- enhanced_visualizer.py:1487-1559 ----------------------------------------------------------
def _xyz(t: torch.Tensor) -> np.ndarray:
    """Device [nz, ny, nx] -> host [nx, ny, nz] (the reference's field order)."""
    return np.ascontiguousarray(t.cpu().numpy().transpose(2, 1, 0))


def export_data_for_analysis(solver, simulation_time: float, step_num: int, particles=None, filter_system=None,
                             scales: Optional[DiagnosticScales] = None, reynolds_number: float = 0.0,
                             out_dir: str = ".") -> list:
    """Write cfd_data_export_step_NNNN.json / .npz with the reference's keys.
    |u| and the pressure rho * c_s^2 are formed on the device
    (pol_export_fields) in the per-element arithmetic of the reference's
    float32 numpy expressions, so the files match what the reference writes
    from the same fields."""
    _ensure_macro(solver)
    ctx = _ctx(solver)
    cfg = ctx.config
    sc = scales or DiagnosticScales.reference(cfg.nz)
    mag = torch.empty_like(ctx.rho)
    pres = torch.empty_like(ctx.rho)
    L.check(ctx.lib.pol_export_fields(ctx.handle, mag.data_ptr(), pres.data_ptr(), L.stream_handle()), ctx.handle,
            "pol_export_fields")
    torch.cuda.synchronize()
    u = ctx.u
    d: dict = {"velocity_field": {"u_x": _xyz(u[0]).tolist(), "u_y": _xyz(u[1]).tolist(), "u_z": _xyz(u[2]).tolist(),
                                  "magnitude": _xyz(mag).tolist()},
               "density_field": _xyz(ctx.rho).tolist(), "pressure_field": _xyz(pres).tolist()}
    if particles is not None:
        act = particles.active.cpu().numpy()
        pos = particles.pos.cpu().numpy().T
        vel = particles.vel.cpu().numpy().T
        d["particles"] = {"positions": pos[act == 1].tolist(), "velocities": vel[act == 1].tolist(),
                          "count": int(np.sum(act))}
    if filter_system is not None and hasattr(filter_system, "get_filter_statistics"):
        d["filter"] = filter_system.get_filter_statistics()
    d["metadata"] = {"simulation_time": simulation_time, "step_number": step_num,
                     "grid_size": [cfg.nx, cfg.ny, cfg.nz],
                     "physical_parameters": {"scale_length": sc.scale_length, "scale_time": sc.scale_time,
                                             "reynolds_number": reynolds_number}}
    base = os.path.join(out_dir, f"cfd_data_export_step_{step_num:04d}")
    with open(base + ".json", "w") as f:
        json.dump(d, f, indent=2)
    np.savez_compressed(base + ".npz", **{k: np.array(v) for k, v in d.items() if isinstance(v, (list, np.ndarray))})
    return [base + ".json", base + ".npz"]
