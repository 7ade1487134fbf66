"""Body-force producers of the main loop on the HIP engine.

Mirrors (reference):
  src/physics/precise_pouring.py:12-330     PrecisePouringSystem
  src/physics/pressure_gradient_drive.py:11-400  PressureGradientDrive
  src/physics/filter_paper.py:616-746       FilterPaperSystem's particle
                                            blocking and dynamic blockage
with the same constants, state and method names.  The per-cell work runs in
csrc/forcing.hip (pol_pour_force / pol_pour_phase / pol_density_drive /
pol_pressure_force / pol_filter_blockage / pol_filter_block_particles) on
the solver context's rho, body force and cell flags; the scalars of the
pour state (time, position of the spiral) are kept on the host in float32,
as the reference keeps them in 0-d float32 fields.  There is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional

import numpy as np
import torch

from . import _lib as L
from .fields import DeviceField

F = np.float32

# the reference's config values (config/core.py, config/physics.py)
NX = NY = NZ = 224
SCALE_LENGTH = 0.14 / 224
SCALE_TIME = (SCALE_LENGTH / 0.01) * 1.2
GRID_SIZE_CM = SCALE_LENGTH * 100
CUP_HEIGHT = 0.085
POUR_RATE_ML_S = 4.0
INLET_DIAMETER_M = 0.005
INLET_AREA = math.pi * (INLET_DIAMETER_M / 2.0) ** 2
CS2 = 1.0 / 3.0


def inlet_velocity() -> float:
    """config/physics.py compute_inlet_velocity: flow / area, +5 %, in lattice units, capped at 0.05."""
    base = POUR_RATE_ML_S * 1e-6 / INLET_AREA
    return min(0.05, base * (1.0 + 0.05) * SCALE_TIME / SCALE_LENGTH)


def _ctx_of(solver):
    return solver.ctx if hasattr(solver, "ctx") else solver


def _tensor(x) -> torch.Tensor:
    return x.to_torch() if isinstance(x, DeviceField) else x


class PrecisePouringSystem:
    """precise_pouring.py:12 on the device.  `solver` supplies the body force
    and cell flags (SolverConfig(force_field=True))."""

    def __init__(self, solver, grid=None):
        self.ctx = _ctx_of(solver)
        self.lib = self.ctx.lib
        cfg = self.ctx.config
        self.nx, self.ny, self.nz = grid or (cfg.nx, cfg.ny, cfg.nz)
        self.POUR_DIAMETER_CM = 0.5
        self.POUR_DIAMETER_GRID = self.POUR_DIAMETER_CM / GRID_SIZE_CM
        self.POUR_VELOCITY = inlet_velocity()
        cup_height_lu = int(CUP_HEIGHT / SCALE_LENGTH)
        v60_top_z = int(5.0 + cup_height_lu)
        self.POUR_HEIGHT = max(8, min(int(v60_top_z + 2), self.nz - 6))
        self.pouring_active = 0
        self.pour_center_x = F(0)
        self.pour_center_y = F(0)
        self.pour_flow_rate = F(0)
        self.pour_pattern = 0
        self.pour_time = F(0)
        self.spiral_radius = F(0)
        self.spiral_speed = F(0)
        self.spiral_center_x = F(0)
        self.spiral_center_y = F(0)

    def start_pouring(self, center_x=None, center_y=None, flow_rate=1.0, pattern="center") -> None:
        center_x = self.nx // 2 if center_x is None else center_x
        center_y = self.ny // 2 if center_y is None else center_y
        self.pour_center_x, self.pour_center_y = F(center_x), F(center_y)
        self.pour_flow_rate = F(flow_rate)
        self.pouring_active = 1
        self.pour_time = F(0)
        if pattern == "center":
            self.pour_pattern = 0
        elif pattern == "spiral":
            self.pour_pattern = 1
            self.spiral_center_x, self.spiral_center_y = F(center_x), F(center_y)
            self.spiral_radius, self.spiral_speed = F(5.0), F(1.0)

    def stop_pouring(self) -> None:
        self.pouring_active = 0

    def adjust_flow_rate(self, rate: float) -> None:
        self.pour_flow_rate = F(rate)

    def _position(self):
        """precise_pouring.py:83-100 (float32)."""
        x, y = self.pour_center_x, self.pour_center_y
        if self.pour_pattern == 1:
            t = F(self.pour_time * self.spiral_speed)
            r = F(self.spiral_radius * (F(1.0) + F(0.1) * t))
            x = F(self.spiral_center_x + r * np.cos(t))
            y = F(self.spiral_center_y + r * np.sin(t))
            d = self.POUR_DIAMETER_GRID
            x = max(F(d), min(F(self.nx - d), x))
            y = max(F(d), min(F(self.ny - d), y))
        return F(x), F(y)

    def _desc(self, dt) -> L.PolPourDesc:
        x, y = self._position()
        d = L.PolPourDesc()
        d.center_x, d.center_y = float(x), float(y)
        d.radius = self.POUR_DIAMETER_GRID / 2.0
        d.stream_height = 4.0
        d.velocity = self.POUR_VELOCITY
        d.flow_rate = float(self.pour_flow_rate)
        d.dt = float(F(dt))
        d.height = int(self.POUR_HEIGHT)
        return d

    def apply_pouring_force(self, *args, dt=None) -> None:
        """precise_pouring.py:132-162, apply_pouring_force(body_force, solid, dt):
        dt is the keyword or the last positional argument (default 1.0); the
        body force and solid mask are the solver context's."""
        if self.pouring_active != 1:
            return
        dt = float(dt) if dt is not None else (float(args[-1]) if args else 1.0)
        self.pour_time = F(self.pour_time + F(dt))
        L.check(self.lib.pol_pour_force(self.ctx.handle, C.byref(self._desc(dt)), L.stream_handle()),
                self.ctx.handle, "pol_pour_force")

    def apply_gradual_phase_change(self, multiphase_phi, solid=None, dt: float = 1.0) -> None:
        """precise_pouring.py:164-193 on the phase field (float32 device [nz, ny, nx])."""
        if self.pouring_active != 1:
            return
        phi = _tensor(multiphase_phi)
        if phi.dtype != torch.float32 or not phi.is_cuda or not phi.is_contiguous():
            raise TypeError("apply_gradual_phase_change needs a contiguous float32 device phase field")
        L.check(self.lib.pol_pour_phase(self.ctx.handle, C.byref(self._desc(dt)), phi.data_ptr(), L.stream_handle()),
                self.ctx.handle, "pol_pour_phase")

    def get_pouring_info(self) -> dict:
        if self.pouring_active != 1:
            return {"active": False, "position": (0, 0), "diameter_grid": 0, "diameter_cm": 0, "velocity": 0,
                    "flow_rate": 0, "pour_time": 0, "pattern": 0}
        x, y = self._position()
        return {"active": True, "position": (float(x), float(y)), "diameter_grid": float(self.POUR_DIAMETER_GRID),
                "diameter_cm": float(self.POUR_DIAMETER_CM), "velocity": float(self.POUR_VELOCITY),
                "flow_rate": float(self.pour_flow_rate), "pour_time": float(self.pour_time),
                "pattern": int(self.pour_pattern)}

    def get_current_flow_rate(self) -> float:
        """Volume flow in lattice units (precise_pouring.py:277-290)."""
        if self.pouring_active == 0:
            return 0.0
        return float(self.pour_flow_rate * inlet_velocity() * INLET_AREA / SCALE_LENGTH ** 2)

    def get_current_flow_rate_ml_s(self) -> float:
        return float(POUR_RATE_ML_S * max(0.0, float(self.pour_flow_rate)))


class PressureGradientDrive:
    """pressure_gradient_drive.py:11 on the device."""

    def __init__(self, lbm_solver):
        self.lbm = lbm_solver
        self.ctx = _ctx_of(lbm_solver)
        self.lib = self.ctx.lib
        cfg = self.ctx.config
        shape = (cfg.nz, cfg.ny, cfg.nx)
        self.HIGH_PRESSURE_RATIO = 1.8
        self.LOW_PRESSURE_RATIO = 0.4
        self.HIGH_PRESSURE_ZONE = 0.8
        self.LOW_PRESSURE_ZONE = 0.2
        self.ADJUSTMENT_RATE = 0.025
        self.MAX_PRESSURE_FORCE = 0.12
        self.GRADIENT_SMOOTHING = 0.1
        self.density_drive_active = 0
        self.force_drive_active = 0
        self.mixed_drive_active = 0
        dev = self.ctx.device
        self._target = torch.zeros(shape, dtype=torch.float32, device=dev)
        self._pforce = torch.zeros((3,) + shape, dtype=torch.float32, device=dev)
        self.initialize_target_density()

    @property
    def target_density(self) -> DeviceField:
        return DeviceField(self._target, "scalar")

    @property
    def pressure_force(self) -> DeviceField:
        return DeviceField(self._pforce, "vector")

    def initialize_target_density(self) -> None:
        """pressure_gradient_drive.py:55-72 (a function of z only, float32)."""
        nz = self._target.shape[0]
        zr = np.arange(nz).astype(F) / F(nz)
        hz, lz = self.HIGH_PRESSURE_ZONE, self.LOW_PRESSURE_ZONE
        nh = (zr - F(hz)) / F(1.0 - hz)
        nd = zr / F(lz)
        t = np.where(zr >= F(hz), F(1.0) + nh * F(self.HIGH_PRESSURE_RATIO - 1.0),
                     np.where(zr <= F(lz), F(self.LOW_PRESSURE_RATIO) + nd * F(1.0 - self.LOW_PRESSURE_RATIO),
                              F(1.0))).astype(F)
        self._target.copy_(torch.from_numpy(t).view(-1, 1, 1).expand_as(self._target))

    def activate_density_drive(self, enable=True):
        self.density_drive_active, self.force_drive_active, self.mixed_drive_active = int(bool(enable)), 0, 0

    def activate_force_drive(self, enable=True):
        self.force_drive_active, self.density_drive_active, self.mixed_drive_active = int(bool(enable)), 0, 0

    def activate_mixed_drive(self, enable=True):
        self.mixed_drive_active, self.density_drive_active, self.force_drive_active = int(bool(enable)), 0, 0

    def _rho_current(self) -> None:
        ens = getattr(self.lbm, "_ensure_macro", None)
        if ens is not None:
            ens()

    def apply_density_drive(self) -> None:
        if self.density_drive_active <= 0:
            return
        self._rho_current()
        L.check(self.lib.pol_density_drive(self.ctx.handle, self._target.data_ptr(), self.ADJUSTMENT_RATE, 0, 0.001,
                                           0.5, 2.0, L.stream_handle()), self.ctx.handle, "pol_density_drive")

    def compute_pressure_gradient(self, mode: int = 0) -> None:
        self._rho_current()
        L.check(self.lib.pol_pressure_force(self.ctx.handle, self._pforce.data_ptr(), int(mode),
                                            self.MAX_PRESSURE_FORCE, CS2, L.stream_handle()),
                self.ctx.handle, "pol_pressure_force")

    def apply_force_drive(self) -> None:
        if self.force_drive_active == 0:
            return
        self.compute_pressure_gradient(mode=1)  # compute and accumulate in one pass

    def apply(self, step: int = 0) -> None:
        """pressure_gradient_drive.py:238-256."""
        if self.density_drive_active == 1:
            self.apply_density_drive()
        if self.force_drive_active == 1:
            self.apply_force_drive()
            return
        if self.mixed_drive_active == 1:
            self.compute_pressure_gradient(mode=2)

    def apply_with_surface_tension(self, multiphase, step: int = 0) -> None:
        """apply(step) followed by multiphase.accumulate_surface_tension_pre_collision()
        (main.py:786-800).  When the drive adds its force to the body force
        (force or mixed drive) and the phase field is bound to the same
        context, both run as one device pass with the same result bit for
        bit; otherwise the two calls run in a row."""
        mode = 1 if self.force_drive_active == 1 else (2 if self.mixed_drive_active == 1 else 0)
        same_ctx = getattr(multiphase, "ctx", None) is self.ctx
        if mode == 0 or not same_ctx or self.density_drive_active == 1:
            self.apply(step)
            multiphase.accumulate_surface_tension_pre_collision()
            return
        self.compute_pressure_gradient(mode=mode | L.POL_PRESSURE_SURFACE_TENSION)

    def get_status(self) -> dict:
        return {"density_drive": bool(self.density_drive_active), "force_drive": bool(self.force_drive_active),
                "mixed_drive": bool(self.mixed_drive_active)}

    def get_statistics(self) -> dict:
        """pressure_gradient_drive.py:281-316 (pressure = rho cs2 over fluid cells)."""
        self._rho_current()
        rho = self.ctx.rho
        flags = getattr(self.ctx, "flags", None)
        fluid = torch.ones_like(rho, dtype=torch.bool) if flags is None else \
            ((flags.view(torch.int32) >> L.POL_KIND_SHIFT) & L.POL_KIND_MASK) != L.POL_KIND_SOLID
        p = rho.double()[fluid] * CS2
        if p.numel() == 0:
            mx, mn, avg = 0.0, 999.0, 0.0
        else:
            mx, mn, avg = max(0.0, float(p.max())), min(999.0, float(p.min())), float(p.mean())
        return {"max_pressure": mx, "min_pressure": mn, "avg_pressure": avg, "pressure_drop": mx - mn,
                "pressure_ratio": mx / mn if mn > 0 else 0}


class FilterPaperSystem:
    """The particle blocking and dynamic blockage of filter_paper.py (the
    fluid resistance of the paper is the porous layer of the fused step)."""

    def __init__(self, lbm_solver, filter_zone: np.ndarray, seed: int = 0):
        """filter_zone: [nz, ny, nx] mask of the paper cells (the V60
        geometry's porous layer)."""
        self.lbm = lbm_solver
        self.ctx = _ctx_of(lbm_solver)
        self.lib = self.ctx.lib
        dev = self.ctx.device
        z = np.ascontiguousarray(np.asarray(filter_zone) != 0).astype(np.uint8)
        self._zone = torch.from_numpy(z).to(dev)
        # the paper cells, listed once: the blockage update touches only them
        self._cells = torch.from_numpy(np.flatnonzero(z.ravel() == 1).astype(np.int32)).to(dev)
        self._acc = torch.zeros(z.shape, dtype=torch.float32, device=dev)
        self._blockage = torch.zeros(z.shape, dtype=torch.float32, device=dev)
        self.seed = int(seed)
        self._calls = 0

    @property
    def filter_zone(self) -> DeviceField:
        return DeviceField(self._zone, "scalar")

    @property
    def accumulated_particles(self) -> DeviceField:
        return DeviceField(self._acc, "scalar")

    @property
    def filter_blockage(self) -> DeviceField:
        return DeviceField(self._blockage, "scalar")

    def update_dynamic_resistance(self) -> None:
        L.check(self.lib.pol_filter_blockage(self.ctx.handle, self._cells.data_ptr(), int(self._cells.numel()),
                                             self._acc.data_ptr(), self._blockage.data_ptr(), L.stream_handle()),
                self.ctx.handle, "pol_filter_blockage")

    def block_particles_at_filter(self, positions, velocities, radii=None, active=None, count=None,
                                  scale: float = SCALE_LENGTH) -> None:
        """positions / velocities: float32 device [3, n]; active int32 [n].
        Grid cell = int(position / scale), as in the reference."""
        pos, vel = _tensor(positions), _tensor(velocities)
        n = int(count if count is not None else pos.shape[1])
        act = _tensor(active) if active is not None else torch.ones(n, dtype=torch.int32, device=pos.device)
        for t, name in ((pos, "positions"), (vel, "velocities")):
            if t.dtype != torch.float32 or not t.is_contiguous() or t.shape[0] != 3:
                raise TypeError("%s must be a contiguous float32 [3, n] device tensor" % name)
        seed = (self.seed * 1000003 + self._calls) & 0xFFFFFFFF
        self._calls += 1
        L.check(self.lib.pol_filter_block_particles(self.ctx.handle, self._zone.data_ptr(), self._acc.data_ptr(),
                                                    pos.data_ptr(), vel.data_ptr(), act.data_ptr(), n, float(scale),
                                                    seed, L.stream_handle()),
                self.ctx.handle, "pol_filter_block_particles")

    def step(self, particle_system: Optional[object] = None) -> None:
        """filter_paper.py:748-790 without apply_filter_effects (the paper's
        flow resistance is applied inside the stream-collide step)."""
        if particle_system is not None:
            self.block_particles_at_filter(particle_system.pos, particle_system.vel, None, particle_system.active)
        self.update_dynamic_resistance()

    def get_filter_statistics(self) -> dict:
        z = self._zone.bool()
        n = int(z.sum().item())
        return {"filter_nodes": n,
                "average_blockage": float(self._blockage[z].double().mean().item()) if n else 0.0,
                "max_blockage": float(self._blockage[z].max().item()) if n else 0.0,
                "total_accumulated_particles": float(self._acc[z].double().sum().item()) if n else 0.0}
