"""Lagrangian coffee particles two-way coupled to the HIP lattice.

Host mirror of CoffeeParticleSystem's coupling (src/physics/coffee_particles.py:
30-64 state, :1101-1105 clear_reaction_forces, :1107-1154
compute_two_way_coupling_forces, :1200-1212 apply_under_relaxation,
:641-720 update_particle_physics, :833-891 get_particle_statistics,
:1214-1245 get_coupling_diagnostics) and of
LBMSolver.step_with_two_way_coupling (src/core/legacy/lbm_solver.py:
1478-1509).  All particle arithmetic runs in the HIP kernels of
csrc/particles.hip through the C ABI; state lives in device tensors (SoA,
[3, P] for vectors) of the lattice's real type.

Two forms of the reaction on the fluid (`reaction=`):
  "implicit" (default): the coupling spreads (w beta v, w beta) into the
      lattice's particle-drag field; the next stream-collide step applies
      sum w beta (v - u) with the cell velocity u solved implicitly, which
      stays stable for any drag coefficient.  The field never needs a
      full-grid clear: advance (or a repeated coupling) zeroes the 8-corner
      footprint it wrote.
  "explicit": the reference's form: the coupling clears reaction_force_field
      and spreads -drag_new into it; add_particle_reaction_forces adds it to
      the body force on the fluid cells before the step.

One coupled step (coupled_step / pol_coupled_step):
  couple   : interpolate u at the particles (u field of the previous step,
             the reference's ordering), Schiller-Naumann drag, relaxation,
             the reaction as above
  LBM step : the fused stream-collide reads the drag field (implicit) or the
             body force (explicit) and writes the new u field
  advance  : zero the drag-field footprint, move the particles
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np
import torch

from . import _lib as L
from .engine import HipContext
from .fields import DeviceField


class HipCoffeeParticleSystem:
    REACTIONS = {"implicit": L.POL_REACTION_IMPLICIT, "explicit": L.POL_REACTION_EXPLICIT}

    def __init__(self, ctx: HipContext, pos: np.ndarray, radius: np.ndarray, mass: np.ndarray, *,
                 rho_fluid: float = 1.0, relax_alpha: float = 0.8, gravity=(0.0, 0.0, 0.0), dt: float = 1.0,
                 max_disp: float = 1.0, wall_damping: float = 0.3, drag_law: int = L.POL_DRAG_TEXTBOOK,
                 drag_max_per_mass: float = 1.0, reaction: str = "implicit"):
        if reaction not in self.REACTIONS:
            raise ValueError("reaction must be 'implicit' or 'explicit'")
        self.reaction = reaction
        if reaction == "implicit" and ctx.pdrag is None:
            raise ValueError("the implicit reaction needs the particle-drag field (SolverConfig.particle_field=True)")
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
        # explicit form: the reference's reaction_force_field, [3, nz, ny, nx]
        self._reaction = (torch.zeros((3,) + tuple(ctx.shape), dtype=rd, device=dev)
                          if reaction == "explicit" else None)
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
        d.reaction = self.REACTIONS[reaction]
        self.desc = d
        self._bind()

    def _bind(self) -> None:
        b = L.PolParticleBuffers()
        for name in ("pos", "vel", "radius", "mass", "active", "drag_new", "drag_old", "drag", "u_fluid", "re_p", "cd"):
            setattr(b, name, getattr(self, name).data_ptr())
        b.drag_field = self.ctx.pdrag.data_ptr() if self.ctx.pdrag is not None else None
        b.u_field = self.ctx.u.data_ptr()
        b.reaction_field = self._reaction.data_ptr() if self._reaction is not None else None
        L.check(self.lib.pol_particles_bind(self.ctx.handle, C.byref(self.desc), C.byref(b)), self.ctx.handle,
                "pol_particles_bind")

    # ---- kernels ------------------------------------------------------------
    def couple(self) -> None:
        L.check(self.lib.pol_particles_couple(self.ctx.handle, L.stream_handle()), self.ctx.handle,
                "pol_particles_couple")

    def advance(self) -> None:
        L.check(self.lib.pol_particles_advance(self.ctx.handle, L.stream_handle()), self.ctx.handle,
                "pol_particles_advance")

    # ---- the reference's coupling verbs (coffee_particles.py) ------------------
    @property
    def reaction_force_field(self) -> DeviceField:
        """coffee_particles.py:52: the reaction field, (NX, NY, NZ, 3) on the host.
        The explicit form holds -drag spread to the cells; the implicit form
        has no such field and presents the particle force the next step
        applies, sum w beta (v - u), formed from the drag field and the
        lattice velocity."""
        if self._reaction is not None:
            return DeviceField(self._reaction, "vector")
        return DeviceField(self.implicit_reaction(), "vector")

    def implicit_reaction(self, u: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Implicit form: per-cell force BV - B u of the drag field with the
        lattice velocity u (default: the context's u buffer), [3, nz, ny, nx]."""
        pd = self.ctx.pdrag
        if pd is None:
            raise ValueError("no implicit drag field on this lattice")
        u = self.ctx.u if u is None else u
        return pd[:3] - pd[3:4] * u

    def compute_two_way_coupling_forces(self, fluid_u=None) -> None:
        """coffee_particles.py:1107-1154: interpolate fluid_u (a DeviceField,
        a [3, nz, ny, nx] device tensor, or None for the lattice's u buffer),
        drag_new, Reynolds number and C_D, and the reaction (the explicit
        form clears reaction_force_field first).  No under-relaxation."""
        uptr = None
        if fluid_u is not None:
            t = fluid_u.to_torch() if isinstance(fluid_u, DeviceField) else fluid_u
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == self.ctx.u.dtype
                    and tuple(t.shape) == tuple(self.ctx.u.shape) and t.is_contiguous()):
                raise ValueError("fluid_u must be a contiguous [3, nz, ny, nx] device tensor of the lattice dtype")
            uptr = t.data_ptr()
        L.check(self.lib.pol_particles_compute_coupling(self.ctx.handle, uptr, L.stream_handle()), self.ctx.handle,
                "pol_particles_compute_coupling")

    def apply_under_relaxation(self, relaxation_factor: float) -> None:
        """coffee_particles.py:1200-1212."""
        L.check(self.lib.pol_particles_relax(self.ctx.handle, float(relaxation_factor), L.stream_handle()),
                self.ctx.handle, "pol_particles_relax")

    def clear_reaction_forces(self) -> None:
        """coffee_particles.py:1101-1105 (explicit: whole field; implicit: the
        footprint of a coupling not yet consumed)."""
        L.check(self.lib.pol_particles_clear_reaction(self.ctx.handle, L.stream_handle()), self.ctx.handle,
                "pol_particles_clear_reaction")

    def update_particle_physics(self, dt: float = 1.0, center_x: float = 0.0, center_y: float = 0.0,
                                bottom_z: float = 0.0, bottom_radius_lu: float = 0.0,
                                top_radius_lu: float = 0.0) -> None:
        """coffee_particles.py:641-720 signature; the push is this build's
        integrator (implicit linearised drag, buoyancy-corrected gravity,
        displacement clamp, rejection at walls and the filter from the cell
        flags, so the cone arguments are accepted and unused)."""
        if abs(float(dt) - self.desc.dt) > 0.0:
            self.desc.dt = float(dt)
            self._bind()
        self.advance()

    def get_particle_statistics(self) -> dict:
        """coffee_particles.py:833-891 keys.  Radii and positions in lattice
        units; a particle counts as valid when its radius is finite and > 0
        and its position finite and inside [0, n] per axis (the reference's
        metre bounds MIN_RADIUS / MAX_RADIUS do not apply to lattice units)."""
        act = self.active.bool().cpu().numpy()
        pos = self.pos.double().cpu().numpy()[:, act].T
        rad = self.radius.double().cpu().numpy()[act]
        nz, ny, nx = self.ctx.shape
        hi = np.array([nx, ny, nz], dtype=np.float64)
        ok = np.isfinite(rad) & (rad > 0) & np.all(np.isfinite(pos), axis=1) & np.all((pos >= 0) & (pos <= hi), axis=1)
        radii, positions = rad[ok], pos[ok]
        n_ok, n_bad = int(ok.sum()), int((~ok).sum())
        coord_err = int((~np.all(np.isfinite(pos), axis=1)).sum())
        stats = (float(radii.mean()), float(radii.std()), float(radii.min()), float(radii.max())) if n_ok else (0.0,) * 4
        return {"count": n_ok, "invalid_count": n_bad, "coordinate_errors": coord_err, "boundary_violations": 0,
                "mean_radius": stats[0], "std_radius": stats[1], "min_radius": stats[2], "max_radius": stats[3],
                "positions": positions, "radii": radii, "success_rate": n_ok / max(1, n_ok + n_bad) * 100}

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
        if self._reaction is not None:  # the reference's figure: max |reaction| over the cells
            rmax = float(torch.linalg.vector_norm(self._reaction.double(), dim=0).max())
        else:
            rmax = float(torch.linalg.vector_norm(self.drag_new.double()[:, act], dim=0).max())
        return {"active_particles": n, "avg_reynolds": float(re.mean()), "max_reynolds": float(re.max()),
                "avg_drag_coeff": float(self.cd[act].double().mean()), "max_reaction_force": rmax,
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
