"""Thermal D3Q7 solver on the HIP engine.

Mirrors (reference):
  src/physics/thermal_lbm.py:40-556     ThermalLBM (same constructor
                                        arguments, relaxation time and its
                                        clamp, method names and return values)
  src/core/thermal_fluid_coupled.py:143-235  the coupled step: fluid step,
                                        velocity hand-over every
                                        coupling_frequency steps, heat source
                                        reset, thermal sub-cycles
The step runs csrc/thermal.hip (pol_thermal_*): one collide-and-push kernel
and one temperature kernel per step, populations and temperatures ping-pong
on the device.  Fields are device tensors in [nz, ny, nx] order; the
DeviceField views (`temperature`, `heat_source`, ...) read and write in the
reference's [x, y, z] order.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np
import torch

from . import _lib as L
from .engine import SolverConfig
from .fields import DeviceField

Q_THERMAL = 7
W_THERMAL = np.array([0.25] + [0.125] * 6, dtype=np.float32)
CS2_THERMAL = 1.0 / 3.0  # the reference's constant for tau (thermal_lbm.py:37); the D3Q7 lattice's own is 1/4


class _BareContext:
    """A pol_ctx for the grid alone (no fluid populations bound)."""

    def __init__(self, nx: int, ny: int, nz: int, precision: str, device: int):
        self.lib = L.load()
        self.config = SolverConfig(nx, ny, nz, precision=precision, periodic=(False, False, False), device=device)
        self.device = torch.device("cuda", device)
        self._h = C.c_void_p()
        L.check(self.lib.pol_create(C.byref(self._h), C.byref(self.config.desc(False))), None, "pol_create")

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            self.lib.pol_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ThermalLBM:
    """thermal_lbm.py:40-556 on the device."""

    def __init__(self, thermal_diffusivity: float = 1.6e-7, scale_length: float = 0.000625,
                 scale_time: float = 0.0625, *, solver=None, grid: Tuple[int, int, int] = (224, 224, 224),
                 precision: str = "f32", device: int = 0):
        if not torch.cuda.is_available():
            raise RuntimeError("ThermalLBM needs a HIP device (there is no CPU fallback)")
        self.alpha_phys = thermal_diffusivity
        self.scale_length = scale_length
        self.scale_time = scale_time
        self.alpha_lu = self.alpha_phys * self.scale_time / (self.scale_length ** 2)
        self.tau_thermal = self.alpha_lu / CS2_THERMAL + 0.5
        if self.tau_thermal < 0.51:
            self.tau_thermal = 0.51  # thermal_lbm.py:79-82
        self.omega_thermal = 1.0 / self.tau_thermal
        if solver is not None:
            self.ctx = solver.ctx if hasattr(solver, "ctx") else solver
            cfg = self.ctx.config
        else:
            self.ctx = _BareContext(*grid, precision, device)
            cfg = self.ctx.config
        self.nx, self.ny, self.nz = cfg.nx, cfg.ny, cfg.nz
        self.lib = self.ctx.lib
        rd = cfg.real_dtype
        dev = self.ctx.device
        shape = (self.nz, self.ny, self.nx)
        self._g = [torch.zeros((Q_THERMAL,) + shape, dtype=rd, device=dev) for _ in range(2)]
        self._T = [torch.zeros(shape, dtype=rd, device=dev) for _ in range(2)]
        self._hs = torch.zeros(shape, dtype=rd, device=dev)
        self._hs_base = torch.zeros(shape, dtype=rd, device=dev)
        self._vel = torch.zeros((3,) + shape, dtype=rd, device=dev)
        self._k = torch.full(shape, 0.68, dtype=rd, device=dev)
        self._cp = torch.full(shape, 4180.0, dtype=rd, device=dev)
        self._alpha = torch.full(shape, self.alpha_phys, dtype=rd, device=dev)
        self._flux = torch.zeros((3,) + shape, dtype=rd, device=dev)
        self.enable_convection = False
        self.reset_source_each_step = False
        b = L.PolThermalBuffers()
        b.g[0], b.g[1] = self._g[0].data_ptr(), self._g[1].data_ptr()
        b.temperature[0], b.temperature[1] = self._T[0].data_ptr(), self._T[1].data_ptr()
        b.heat_source = self._hs.data_ptr()
        b.velocity = self._vel.data_ptr()
        b.conductivity = self._k.data_ptr()
        b.heat_source_base = self._hs_base.data_ptr()
        L.check(self.lib.pol_thermal_bind(self.ctx.handle, C.byref(self._desc()), C.byref(b)), self.ctx.handle,
                "pol_thermal_bind")
        self._gcur = 0
        self._tcur = 0
        self.is_initialized = False
        self.current_step = 0
        self._last_unstable = 0

    def _desc(self) -> L.PolThermalDesc:
        d = L.PolThermalDesc()
        d.omega = self.omega_thermal
        d.options = (L.POL_THERMAL_CONVECTION if self.enable_convection else 0) | \
            (L.POL_THERMAL_RESET_SOURCE if self.reset_source_each_step else 0)
        return d

    # ---- fields (reference index order [x, y, z]) ------------------------------------
    @property
    def temperature(self) -> DeviceField:
        return DeviceField(self._T[self._tcur], "scalar")

    @property
    def temperature_old(self) -> DeviceField:
        return DeviceField(self._T[1 - self._tcur], "scalar")

    @property
    def heat_source(self) -> DeviceField:
        return DeviceField(self._hs, "scalar")

    @property
    def velocity_field(self) -> DeviceField:
        return DeviceField(self._vel, "vector")

    @property
    def heat_flux(self) -> DeviceField:
        """-k grad T of the current temperature, formed when read (the
        reference forms it every step, thermal_lbm.py:421; nothing in the
        step reads it)."""
        return DeviceField(self._flux, "vector", before_read=self.compute_heat_flux)

    @property
    def thermal_conductivity(self) -> DeviceField:
        return DeviceField(self._k, "scalar")

    @property
    def heat_capacity(self) -> DeviceField:
        return DeviceField(self._cp, "scalar")

    @property
    def thermal_diffusivity_field(self) -> DeviceField:
        return DeviceField(self._alpha, "scalar")

    @property
    def g(self) -> np.ndarray:
        """Current populations in the reference's [nx, ny, nz, 7] order (host copy)."""
        return np.ascontiguousarray(self._g[self._gcur].cpu().numpy().transpose(3, 2, 1, 0))

    @property
    def g_device(self) -> torch.Tensor:
        """Current populations on the device, [7, nz, ny, nx]."""
        return self._g[self._gcur]

    # ---- initialisation --------------------------------------------------------------
    def init_temperature_field(self, T_initial: float, T_hot_region: float, hot_region_height: int) -> None:
        """thermal_lbm.py:132-170: T = T_hot below hot_region_height, else T_initial; g = w T."""
        T = self._T[self._tcur]
        z = torch.arange(self.nz, device=T.device).view(-1, 1, 1)
        T.copy_(torch.where(z < hot_region_height, torch.tensor(T_hot_region, dtype=T.dtype, device=T.device),
                            torch.tensor(T_initial, dtype=T.dtype, device=T.device)).expand_as(T))
        self._k.fill_(0.68)
        self._cp.fill_(4180.0)
        self._alpha.fill_(self.alpha_phys)
        self._hs.zero_()
        self._init_populations()

    def _init_populations(self) -> None:
        L.check(self.lib.pol_thermal_init(self.ctx.handle, L.stream_handle()), self.ctx.handle, "pol_thermal_init")

    def complete_initialization(self, T_initial: float, T_hot_region: float, hot_region_height: int) -> None:
        self.init_temperature_field(T_initial, T_hot_region, hot_region_height)
        self.is_initialized = True

    def initialize_from(self, T: np.ndarray) -> None:
        """Start from a temperature field [nx, ny, nz] (g = w T)."""
        self.temperature.from_numpy(np.asarray(T))
        self._init_populations()
        self.is_initialized = True

    # ---- step ------------------------------------------------------------------------
    def step(self) -> bool:
        """thermal_lbm.py:396-427: True when the stability test passes."""
        if not self.is_initialized:
            return False
        u = C.c_int32()
        L.check(self.lib.pol_thermal_step(self.ctx.handle, 1, C.byref(u), L.stream_handle()), self.ctx.handle,
                "pol_thermal_step")
        self._gcur ^= 1
        self._tcur ^= 1
        self._last_unstable = int(u.value)
        if self._last_unstable:
            return False
        self.current_step += 1
        return True

    def run(self, nsteps: int) -> bool:
        """nsteps steps with one stability read-back at the end."""
        if not self.is_initialized:
            return False
        u = C.c_int32()
        L.check(self.lib.pol_thermal_step(self.ctx.handle, nsteps, C.byref(u), L.stream_handle()), self.ctx.handle,
                "pol_thermal_step")
        if nsteps % 2:
            self._gcur ^= 1
            self._tcur ^= 1
        self._last_unstable = int(u.value)
        self.current_step += nsteps
        return self._last_unstable == 0

    def collision_step(self) -> None:
        """Collision and streaming form one fused pass here; see streaming_step."""

    def streaming_step(self) -> None:
        """Runs the fused collide-and-stream pass of one step."""
        self.step()

    def compute_temperature(self) -> None:
        """The temperature is formed by every step (thermal_lbm.py:255-266)."""

    def compute_heat_flux(self) -> None:
        """thermal_lbm.py:268-286 into `heat_flux`."""
        L.check(self.lib.pol_thermal_heat_flux(self.ctx.handle, self._flux.data_ptr(), L.stream_handle()),
                self.ctx.handle, "pol_thermal_heat_flux")

    # ---- boundary forms (thermal_lbm.py:288-346) -------------------------------------------
    def _mask(self, boundary_mask) -> torch.Tensor:
        if isinstance(boundary_mask, DeviceField):
            t = boundary_mask.to_torch()
        elif isinstance(boundary_mask, torch.Tensor):
            t = boundary_mask
        else:
            t = torch.from_numpy(np.ascontiguousarray(np.asarray(boundary_mask).transpose(2, 1, 0)))
        return (t != 0).to(torch.uint8).to(self._hs.device).contiguous()

    def _boundary(self, kind: int, mask, a: float, b: float = 0.0) -> None:
        m = self._mask(mask)
        L.check(self.lib.pol_thermal_boundary(self.ctx.handle, kind, m.data_ptr(), float(a), float(b),
                                              L.stream_handle()), self.ctx.handle, "pol_thermal_boundary")
        torch.cuda.synchronize()

    def apply_dirichlet_bc(self, boundary_mask, boundary_temp: float) -> None:
        self._boundary(1, boundary_mask, boundary_temp)

    def apply_neumann_bc(self, boundary_mask, boundary_flux: float) -> None:
        self._boundary(2, boundary_mask, boundary_flux)

    def apply_convective_bc(self, boundary_mask, h_conv: float, T_ambient: float) -> None:
        self._boundary(3, boundary_mask, h_conv, T_ambient)

    # ---- diagnostics ----------------------------------------------------------------
    def check_numerical_stability(self) -> int:
        """The stability test of the last step (thermal_lbm.py:348-376): 1 unstable, 0 stable."""
        return 1 if self._last_unstable else 0

    def compute_diagnostics(self) -> Tuple[float, float, float]:
        out = np.zeros(3, dtype=np.float64)
        L.check(self.lib.pol_thermal_stats(self.ctx.handle, out.ctypes.data, L.stream_handle()), self.ctx.handle,
                "pol_thermal_stats")
        return float(out[0]), float(out[1]), float(out[2] / (self.nx * self.ny * self.nz))

    def get_temperature_stats(self) -> Tuple[float, float, float]:
        """(min, max, mean) temperature (thermal_lbm.py:429-440)."""
        return self.compute_diagnostics()

    def get_effective_thermal_diffusivity(self) -> float:
        """thermal_lbm.py:442-449 (the reference's formula, c_s^2 = 1/3)."""
        return (self.tau_thermal - 0.5) * CS2_THERMAL * (self.scale_length ** 2) / self.scale_time

    def set_heat_source(self, source_field: np.ndarray) -> None:
        if tuple(source_field.shape) != (self.nx, self.ny, self.nz):
            raise ValueError(f"heat source shape {source_field.shape} != {(self.nx, self.ny, self.nz)}")
        self.heat_source.from_numpy(source_field.astype(np.float32))

    def reset(self) -> None:
        self.current_step = 0
        self.is_initialized = False
        for t in self._g:
            t.zero_()
        self._T[self._tcur].fill_(25.0)
        self._hs.zero_()
        self._vel.zero_()

    # ---- convection coupling (thermal_lbm.py:481-556) ---------------------------------------
    def enable_convection_coupling(self, enable: bool = True) -> None:
        self.enable_convection = bool(enable)
        L.check(self.lib.pol_thermal_set(self.ctx.handle, C.byref(self._desc())), self.ctx.handle, "pol_thermal_set")

    def set_velocity_field(self, velocity_field) -> None:
        """Copy the fluid velocity (a DeviceField / [3, nz, ny, nx] tensor /
        solver with `u`) into the thermal solver's velocity buffer."""
        if not self.enable_convection:
            return
        src = velocity_field
        if hasattr(src, "u") and not isinstance(src, (torch.Tensor, DeviceField)):
            src = src.u
        if isinstance(src, DeviceField):
            src = src.to_torch()
        self._vel.copy_(src)

    def use_base_heat_source(self, base_heat_source, enable: bool = True) -> None:
        """Start every following step from base_heat_source (the coupled
        loop's reset_heat_source_to_base before each single thermal step,
        done inside the step instead of as a separate copy)."""
        if enable:
            src = base_heat_source.to_torch() if isinstance(base_heat_source, DeviceField) else base_heat_source
            if isinstance(src, np.ndarray):
                src = torch.from_numpy(np.ascontiguousarray(src.astype(np.float32).transpose(2, 1, 0)))
            self._hs_base.copy_(src)
        self.reset_source_each_step = bool(enable)
        L.check(self.lib.pol_thermal_set(self.ctx.handle, C.byref(self._desc())), self.ctx.handle, "pol_thermal_set")

    def reset_heat_source_to_base(self, base_heat_source) -> None:
        src = base_heat_source.to_torch() if isinstance(base_heat_source, DeviceField) else base_heat_source
        if isinstance(src, np.ndarray):
            self.heat_source.from_numpy(src.astype(np.float32))
        else:
            self._hs.copy_(src)


@dataclass
class CouplingConfig:
    """thermal_fluid_coupled.py:35-42."""

    coupling_frequency: int = 1
    velocity_smoothing: bool = True
    thermal_subcycles: int = 1
    enable_diagnostics: bool = True
    max_coupling_error: float = 1e6


class ThermalFluidCoupledSolver:
    """thermal_fluid_coupled.py:44-460: a fluid solver and a ThermalLBM that
    shares its context; each step runs the fluid step, hands the velocity
    over every coupling_frequency steps, resets the heat source to its base
    and runs the thermal sub-cycles (:143-209).  With one sub-cycle the reset
    runs inside the thermal step (POL_THERMAL_RESET_SOURCE) instead of as a
    copy.  The reference builds its own 224^3 solvers; here they may be
    passed in (fluid_solver / thermal_solver) or are built on `grid`."""

    def __init__(self, coupling_config: Optional[CouplingConfig] = None, thermal_diffusivity: float = 1.6e-7, *,
                 fluid_solver=None, thermal_solver: Optional[ThermalLBM] = None, grid=(224, 224, 224),
                 precision: str = "f32"):
        self.coupling_config = coupling_config or CouplingConfig()
        if fluid_solver is None:
            from .solver import HipLBMSolver

            fluid_solver = HipLBMSolver(SolverConfig(*grid, precision=precision, periodic=(False, False, False)))
        self.fluid_solver = fluid_solver
        self.thermal_solver = thermal_solver if thermal_solver is not None else \
            ThermalLBM(thermal_diffusivity, solver=fluid_solver)
        self.thermal_solver.enable_convection_coupling(True)
        if hasattr(fluid_solver, "enable_thermal_coupling_output"):
            fluid_solver.enable_thermal_coupling_output(True)
        self.base_heat_source = torch.zeros_like(self.thermal_solver._hs)
        self.fused_reset = self.coupling_config.thermal_subcycles == 1
        if self.fused_reset:
            self.thermal_solver.use_base_heat_source(self.base_heat_source)
        self.coupling_step = 0
        self.is_initialized = False
        self._reset_stats()

    def _reset_stats(self) -> None:
        self.performance_stats = {"fluid_time": 0.0, "thermal_time": 0.0, "coupling_time": 0.0, "total_steps": 0}

    def set_base_heat_source(self, base_heat_source: Optional[np.ndarray]) -> None:
        """The base source [nx, ny, nz] every step starts from (zeros if None)."""
        if base_heat_source is None:
            self.base_heat_source.zero_()
        else:
            b = np.asarray(base_heat_source)
            if b.shape != (self.thermal_solver.nx, self.thermal_solver.ny, self.thermal_solver.nz):
                raise ValueError("heat source shape %s does not match the grid" % (b.shape,))
            self.base_heat_source.copy_(torch.from_numpy(np.ascontiguousarray(b.astype(np.float32).transpose(2, 1, 0))))
            self.thermal_solver.reset_heat_source_to_base(self.base_heat_source)
        if self.fused_reset:
            self.thermal_solver.use_base_heat_source(self.base_heat_source)

    def initialize_system(self, fluid_initial_conditions: dict, thermal_initial_conditions: dict,
                          base_heat_source: Optional[np.ndarray] = None) -> None:
        """thermal_fluid_coupled.py:100-141."""
        self.fluid_solver.init_fields()
        self.thermal_solver.complete_initialization(
            T_initial=thermal_initial_conditions.get("T_initial", 25.0),
            T_hot_region=thermal_initial_conditions.get("T_hot_region", 93.0),
            hot_region_height=thermal_initial_conditions.get("hot_region_height", 20))
        self.set_base_heat_source(base_heat_source)
        self.is_initialized = True

    def step(self) -> bool:
        """thermal_fluid_coupled.py:143-209."""
        if not self.is_initialized:
            return False
        cfg = self.coupling_config
        t0 = time.time()
        self.fluid_solver.step()
        t1 = time.time()
        self.performance_stats["fluid_time"] += t1 - t0
        if self.coupling_step % cfg.coupling_frequency == 0:
            get = getattr(self.fluid_solver, "get_velocity_field_for_thermal_coupling", None)
            self.thermal_solver.set_velocity_field(get() if get is not None else self.fluid_solver.u)
            self.performance_stats["coupling_time"] += time.time() - t1
        t2 = time.time()
        if not self.fused_reset:
            self.thermal_solver.reset_heat_source_to_base(self.base_heat_source)
        for _ in range(cfg.thermal_subcycles):
            if not self.thermal_solver.step():
                return False
        self.performance_stats["thermal_time"] += time.time() - t2
        if cfg.enable_diagnostics and not self._check_coupling_stability():
            return False
        self.coupling_step += 1
        self.performance_stats["total_steps"] += 1
        return True

    def _check_coupling_stability(self) -> bool:
        """thermal_fluid_coupled.py:237-262 (the velocity bound on the device)."""
        tmin, tmax, _ = self.thermal_solver.get_temperature_stats()
        if tmax > 150.0 or tmin < -10.0:
            return False
        if abs(tmax - tmin) > self.coupling_config.max_coupling_error:
            return False
        u = self.thermal_solver._vel
        return not bool((torch.linalg.vector_norm(u, dim=0) > 1.0).any())

    def get_coupling_diagnostics(self) -> dict:
        """thermal_fluid_coupled.py:264-304."""
        ps = self.performance_stats
        if ps["total_steps"] == 0:
            return {"status": "not_started"}
        tmin, tmax, tavg = self.thermal_solver.get_temperature_stats()
        total = ps["fluid_time"] + ps["thermal_time"] + ps["coupling_time"]
        frac = (lambda v: v / total if total > 0 else 0)
        return {"coupling_step": self.coupling_step,
                "thermal_stats": {"T_min": float(tmin), "T_max": float(tmax), "T_avg": float(tavg),
                                  "thermal_diffusivity": self.thermal_solver.get_effective_thermal_diffusivity()},
                "performance": {"fluid_fraction": frac(ps["fluid_time"]), "thermal_fraction": frac(ps["thermal_time"]),
                                "coupling_fraction": frac(ps["coupling_time"]),
                                "steps_per_second": ps["total_steps"] / total if total > 0 else 0},
                "coupling_config": {"frequency": self.coupling_config.coupling_frequency,
                                    "thermal_subcycles": self.coupling_config.thermal_subcycles,
                                    "convection_enabled": self.thermal_solver.enable_convection}}

    def save_coupling_state(self, step_num: int) -> dict:
        """thermal_fluid_coupled.py:306-333 (host copies, reference index order)."""
        return {"step": step_num, "velocity": self.fluid_solver.u.to_numpy(), "density": self.fluid_solver.rho.to_numpy(),
                "temperature": self.thermal_solver.temperature.to_numpy(),
                "heat_flux": self.thermal_solver.heat_flux.to_numpy()}

    def reset_coupling_system(self) -> None:
        """thermal_fluid_coupled.py:335-355."""
        if hasattr(self.fluid_solver, "init_fields"):
            self.fluid_solver.init_fields()
        self.thermal_solver.reset()
        self.coupling_step = 0
        self.is_initialized = False
        self._reset_stats()

    # fluid-solver fields seen through the coupled solver (thermal_fluid_coupled.py:357-415)
    @property
    def solid(self):
        return self.fluid_solver.solid

    @property
    def rho(self):
        return self.fluid_solver.rho

    @property
    def phase(self):
        return getattr(self.fluid_solver, "phase", None)

    @property
    def u(self):
        return self.fluid_solver.u

    def get_temperature_field(self):
        return self.thermal_solver.temperature
