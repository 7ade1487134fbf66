"""Lagrangian coffee particles two-way coupled to the HIP lattice.

Host mirror of the coupling part of CoffeeParticleSystem
(src/physics/coffee_particles.py:30-64 state, :1107-1154
compute_two_way_coupling_forces, :1200-1212 apply_under_relaxation,
:641-720 update_particle_physics) and of LBMSolver.step_with_two_way_coupling
(src/core/legacy/lbm_solver.py:1485-1509).  All particle arithmetic runs in
the HIP kernels of csrc/particles.hip through the C ABI; state lives in
device tensors (SoA, [3, P] for vectors) of the lattice's real type.

One coupled step:
  couple   : interpolate u at the particles (u field of the previous step,
             the reference's ordering), Schiller-Naumann drag, relaxation,
             reaction -drag scattered into the lattice force field
  LBM step : the fused stream-collide reads that force field (Guo forcing)
             and writes the new u field
  advance  : zero the reaction footprint, move the particles
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np
import torch

from . import _lib as L
from .engine import HipContext


class HipCoffeeParticleSystem:
    def __init__(self, ctx: HipContext, pos: np.ndarray, radius: np.ndarray, mass: np.ndarray, *,
                 rho_fluid: float = 1.0, relax_alpha: float = 0.8, gravity=(0.0, 0.0, 0.0), dt: float = 1.0,
                 max_disp: float = 1.0, wall_damping: float = 0.3, drag_law: int = L.POL_DRAG_TEXTBOOK,
                 drag_max_per_mass: float = 1.0):
        if ctx.pdrag is None:
            raise ValueError("the lattice needs the particle-drag field (SolverConfig.particle_field=True)")
        self.ctx = ctx
        self.lib = ctx.lib
        rd = ctx.config.real_dtype
        dev = ctx.device
        P = int(pos.shape[1])
        self.count = P

        def t(a, dtype=rd):
            return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(dev)

        self.pos = t(pos)
        self.vel = torch.zeros((3, P), dtype=rd, device=dev)
        self.radius = t(radius)
        self.mass = t(mass)
        self.active = torch.ones(P, dtype=torch.int32, device=dev)
        self.drag_new = torch.zeros((3, P), dtype=rd, device=dev)
        self.drag_old = torch.zeros((3, P), dtype=rd, device=dev)
        self.drag = torch.zeros((3, P), dtype=rd, device=dev)
        self.u_fluid = torch.zeros((3, P), dtype=rd, device=dev)
        self.re_p = torch.zeros(P, dtype=rd, device=dev)
        self.cd = torch.zeros(P, dtype=rd, device=dev)
        d = L.PolParticleDesc()
        d.count = P
        d.drag_law = int(drag_law)
        d.rho_fluid = rho_fluid
        d.nu_fluid = (ctx.config.tau - 0.5) / 3.0
        d.drag_max_per_mass = drag_max_per_mass
        d.relax_alpha = relax_alpha
        d.gravity = (C.c_double * 3)(*gravity)
        d.dt = dt
        d.max_disp = max_disp
        d.wall_damping = wall_damping
        self.desc = d
        self._bind()

    def _bind(self) -> None:
        b = L.PolParticleBuffers()
        for name in ("pos", "vel", "radius", "mass", "active", "drag_new", "drag_old", "drag", "u_fluid", "re_p", "cd"):
            setattr(b, name, getattr(self, name).data_ptr())
        b.drag_field = self.ctx.pdrag.data_ptr()
        b.u_field = self.ctx.u.data_ptr()
        L.check(self.lib.pol_particles_bind(self.ctx.handle, C.byref(self.desc), C.byref(b)), self.ctx.handle,
                "pol_particles_bind")

    # ---- kernels ------------------------------------------------------------
    def couple(self) -> None:
        L.check(self.lib.pol_particles_couple(self.ctx.handle, L.stream_handle()), self.ctx.handle,
                "pol_particles_couple")

    def advance(self) -> None:
        L.check(self.lib.pol_particles_advance(self.ctx.handle, L.stream_handle()), self.ctx.handle,
                "pol_particles_advance")

    def prime(self) -> None:
        """Fill the lattice u field from the current populations (before the first couple)."""
        self.ctx.macroscopic(macro=True)

    def coupled_step(self, solver=None, relaxation_factor: Optional[float] = None, nsteps: int = 1) -> None:
        """couple -> fused LBM step (writes rho/u) -> advance, nsteps times, in
        one native loop (pol_coupled_step)."""
        if relaxation_factor is not None and abs(relaxation_factor - self.desc.relax_alpha) > 0:
            self.desc.relax_alpha = float(relaxation_factor)
            self._bind()
        st = self.lib.pol_coupled_step(self.ctx.handle, int(nsteps), 0, None, L.stream_handle())
        L.check(st, self.ctx.handle, "pol_coupled_step")
        if solver is not None and hasattr(solver, "_macro_valid"):
            solver._step_count += nsteps
            solver._macro_valid = False

    # ---- host views (reference names) -------------------------------------------
    def positions(self) -> np.ndarray:
        return self.pos.double().cpu().numpy()

    def get_coupling_diagnostics(self) -> dict:
        """coffee_particles.py:1214-1245 keys, reduced on the device."""
        act = self.active.bool()
        n = int(act.sum().item())
        if n == 0:
            return {"active_particles": 0, "avg_reynolds": 0.0, "avg_drag_coeff": 0.0,
                    "max_reaction_force": 0.0, "coupling_quality": "no_particles"}
        re = self.re_p[act].double()
        dmag = torch.linalg.vector_norm(self.drag_new.double()[:, act], dim=0)
        return {"active_particles": n, "avg_reynolds": float(re.mean()), "max_reynolds": float(re.max()),
                "avg_drag_coeff": float(self.cd[act].double().mean()), "max_reaction_force": float(dmag.max()),
                "coupling_quality": "active"}

    def drag_field_sums(self) -> tuple:
        """FP64 sums of the drag field after couple() -- (sum_c BV_c, sum_c B_c) -- and
        the same quantities summed over the particles (sum_p beta_p v_p, sum_p beta_p)."""
        fld = self.ctx.pdrag.double().sum(dim=(1, 2, 3)).cpu().numpy()
        act = self.active.bool()
        rel = self.u_fluid.double() - self.vel.double()
        mag = torch.linalg.vector_norm(rel, dim=0)
        beta = torch.where(mag > 1e-8, torch.linalg.vector_norm(self.drag_new.double(), dim=0) / mag.clamp_min(1e-30),
                           torch.zeros_like(mag))
        beta = torch.where(act, beta, torch.zeros_like(beta))
        bv = (beta[None, :] * self.vel.double()).sum(dim=1).cpu().numpy()
        return fld, np.concatenate([bv, [float(beta.sum())]])


def particle_gravity(g: float, density_ratio: float) -> tuple:
    """Buoyancy-corrected acceleration (0, 0, -g (1 - rho_f / rho_p))."""
    return (0.0, 0.0, -g * (1.0 - 1.0 / density_ratio))
