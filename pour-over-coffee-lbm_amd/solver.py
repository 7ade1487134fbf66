"""HipLBMSolver: the reference's LBM solver interface on the HIP engine.

Implements the method and attribute set of LBMSolverProtocol
(src/core/lbm_protocol.py:12-193) plus the de-facto extras callers use
(main.py:378-442, thermal_fluid_coupled.py:83-221,
numerical_stability.py:298, legacy/lbm_solver.py:656-660,1195-1206,1478-1509).
Collision, streaming and boundaries are one fused launch; the separate
protocol verbs map onto it as documented per method.
"""
from __future__ import annotations

from typing import Any, Dict, Optional, Protocol, Tuple, runtime_checkable

import numpy as np
import torch

from . import _lib as L
from .backends import ComputeBackend, HipBackend, create_backend
from .engine import HipContext, SolverConfig
from .fields import DeviceField
from .lattice import OPP


@runtime_checkable
class LBMSolverProtocol(Protocol):
    """Protocol of src/core/lbm_protocol.py:12-193 (field types left open)."""

    f: Any
    f_new: Any
    rho: Any
    u: Any
    solid: Any
    phase: Any

    def step(self) -> None: ...
    def collision_step(self) -> None: ...
    def streaming_step(self) -> None: ...
    def compute_macroscopic_quantities(self) -> None: ...
    def apply_boundary_conditions(self) -> None: ...
    def initialize_fields(self, initial_density: float = 1.0,
                          initial_velocity: Tuple[float, float, float] = (0.0, 0.0, 0.0)) -> None: ...
    def set_geometry(self, geometry_function: Any) -> None: ...
    def get_diagnostics(self) -> dict: ...
    def check_stability(self) -> bool: ...
    def get_kinetic_energy(self) -> float: ...
    def get_mass_conservation_error(self) -> float: ...
    def get_memory_usage(self) -> dict: ...
    def optimize_memory_layout(self) -> None: ...
    def enable_les_turbulence(self, smagorinsky_constant: float = 0.1) -> None: ...
    def add_force_term(self, force_field: Any) -> None: ...
    def export_vtk(self, filename: str) -> None: ...


@runtime_checkable
class ParticleCoupledLBMProtocol(LBMSolverProtocol, Protocol):
    """src/core/lbm_protocol.py:221-238."""

    def update_particle_coupling(self, particle_system: Any) -> None: ...
    def compute_drag_force(self) -> Any: ...


def validate_lbm_solver(solver: Any) -> bool:
    """isinstance check of src/core/lbm_protocol.py:241-257."""
    return isinstance(solver, LBMSolverProtocol)


class HipMemoryAdapter:
    """Memory adapter carrying the HIP context (attribute contract of cpu_backend.py:207-212)."""

    layout_type = "SoA f[q][z][y][x]"

    def __init__(self, solver: "HipLBMSolver"):
        self.hip_context = _HipContextAdapter(solver)
        self._solver = solver

    @property
    def f(self):
        return self._solver.f

    @property
    def f_new(self):
        return self._solver.f_new

    @property
    def rho(self):
        return self._solver.rho

    @property
    def u(self):
        return self._solver.u

    @property
    def solid(self):
        return self._solver.solid


class _HipContextAdapter:
    """HipContext wrapper that marks solver macros stale after steps."""

    def __init__(self, solver: "HipLBMSolver"):
        self._s = solver
        self.cells = solver.config.cells

    @property
    def config(self) -> SolverConfig:
        return self._s.config

    def step(self, n: int = 1) -> None:
        self._s._advance(n)

    def macroscopic(self) -> None:
        self._s._ensure_macro()

    def synchronize(self) -> None:
        self._s.ctx.synchronize()


class HipLBMSolver:
    """D3Q19 solver on MI355X.

    Parameters come in a SolverConfig; geometry (cell flags) is optional
    (None = all-fluid box with the configured periodicity).
    """

    def __init__(self, config: SolverConfig, flags: Optional[np.ndarray] = None, backend: Optional[ComputeBackend] = None):
        self.config = config
        self.backend = backend if backend is not None else HipBackend()
        if flags is None and not all(config.periodic):
            from .geometry import link_flags

            kind = np.zeros((config.nz, config.ny, config.nx), dtype=np.uint8)
            flags, _ = link_flags(kind, periodic=tuple(bool(p) for p in config.periodic))
        self.ctx = HipContext(config, flags)
        self._flags_host = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint32)
        self.use_les = config.les
        self._macro_valid = False
        self._step_count = 0
        self._mass0: Optional[float] = None
        dev16 = True  # device buffers hold f - w in every precision
        # reference-style field views
        self.rho = DeviceField(self.ctx.rho, "scalar", before_read=self._ensure_macro)
        self.u = DeviceField(self.ctx.u, "vector", before_read=self._ensure_macro)
        self._f_views = [DeviceField(t, "population", after_write=self._invalidate, deviation=dev16, ctx=self.ctx)
                         for t in self.ctx.f]
        self.body_force = DeviceField(self.ctx.force, "vector") if self.ctx.force is not None else None
        self.phase = DeviceField(torch.ones(self.ctx.shape, dtype=config.real_dtype, device=self.ctx.device),
                                 "scalar")
        self._solid_t = torch.zeros(self.ctx.shape, dtype=torch.uint8, device=self.ctx.device)
        if self._flags_host is not None:
            kind = (self._flags_host >> L.POL_KIND_SHIFT) & L.POL_KIND_MASK
            self._solid_t.copy_(torch.from_numpy((kind == L.POL_KIND_SOLID).astype(np.uint8)))
        self.solid = DeviceField(self._solid_t, "scalar")
        self.opposite_dir = OPP.copy()
        self.memory_adapter = HipMemoryAdapter(self)
        self.boundary_manager = None  # boundaries are applied inside the fused step
        self.thermal_coupling_enabled = False
        self._pending_coupling = None  # particle system coupled by update_particle_coupling, not yet stepped
        self._last_particles = None
        # the reference's finite-difference LES (legacy/lbm_solver.py:93-111
        # hands the solver's phase field to the model)
        self.les_model = None
        if config.les_fd:
            from .les import HipLESTurbulenceModel

            self.les_model = HipLESTurbulenceModel(self.ctx)
            self.les_model.set_phase_field(self.phase)
            self.les_nu_sgs = self.les_model.nu_sgs

    # ---- fields ------------------------------------------------------------------
    @property
    def f(self) -> DeviceField:
        """Current populations (post-collision state of the last step)."""
        return self._f_views[self.ctx.current]

    @property
    def f_new(self) -> DeviceField:
        return self._f_views[1 - self.ctx.current]

    @property
    def ux(self) -> DeviceField:
        self._ensure_macro()
        return DeviceField(self.ctx.u[0], "scalar")

    @property
    def uy(self) -> DeviceField:
        self._ensure_macro()
        return DeviceField(self.ctx.u[1], "scalar")

    @property
    def uz(self) -> DeviceField:
        self._ensure_macro()
        return DeviceField(self.ctx.u[2], "scalar")

    @property
    def les_mask(self) -> DeviceField:
        """1 where LES is allowed (filter_paper.py:199-204 turns it off in the paper)."""
        if self._flags_host is None:
            m = np.ones(self.ctx.shape, dtype=np.int32)
        else:
            m = ((self._flags_host & L.POL_FLAG_NO_LES) == 0).astype(np.int32)
        return DeviceField(torch.from_numpy(m).to(self.ctx.device), "scalar")

    @property
    def step_count(self) -> int:
        return self._step_count

    def _invalidate(self) -> None:
        self._macro_valid = False

    def _ensure_macro(self) -> None:
        if not self._macro_valid:
            self.ctx.macroscopic(macro=True)
            self._macro_valid = True

    def _advance(self, n: int) -> None:
        self.ctx.step(n)
        self._step_count += n
        self._macro_valid = False

    # ---- core operations ----------------------------------------------------------
    def step(self) -> None:
        """One fused stream-collide step through the backend (lbm_unified.py:212-222).
        A coupling left by update_particle_coupling acts on this step only."""
        self.backend.execute_collision_streaming(self.memory_adapter, {"tau": self.config.tau, "sync": False})
        self._consume_coupling()

    def run(self, nsteps: int) -> None:
        """nsteps fused steps in one C call (no per-step Python work)."""
        if nsteps <= 0:
            return
        if self._pending_coupling is not None:
            self.step()
            nsteps -= 1
            if nsteps == 0:
                return
        self.backend.execute_collision_streaming(self.memory_adapter,
                                                 {"tau": self.config.tau, "steps": int(nsteps), "sync": False})

    def _consume_coupling(self) -> None:
        ps = self._pending_coupling
        if ps is None:
            return
        self._pending_coupling = None
        self._last_particles = ps
        if getattr(ps, "reaction", "implicit") != "explicit":
            ps.clear_reaction_forces()  # the drag-field footprint of this step's coupling

    def collision_step(self) -> None:
        """Collision is fused with streaming: this runs one full step."""
        self.step()

    def streaming_step(self) -> None:
        """Streaming is fused into collision_step/step; nothing separate to run."""

    def compute_macroscopic_quantities(self) -> None:
        self._macro_valid = False
        self._ensure_macro()

    def apply_boundary_conditions(self) -> None:
        """Walls (halfway bounce-back), inlet and outlet act inside the fused step."""
        self.backend.apply_boundary_conditions(self.memory_adapter)

    def swap_fields(self) -> None:
        """Buffers swap by pointer inside pol_step (legacy/lbm_solver.py:630-654 copied them)."""

    # ---- initialisation -----------------------------------------------------------
    def initialize_fields(self, initial_density: float = 1.0,
                          initial_velocity: Tuple[float, float, float] = (0.0, 0.0, 0.0)) -> None:
        rd = self.config.real_dtype
        rho = torch.full(self.ctx.shape, float(initial_density), dtype=rd, device=self.ctx.device)
        u = torch.empty((3,) + self.ctx.shape, dtype=rd, device=self.ctx.device)
        for i in range(3):
            u[i].fill_(float(initial_velocity[i]))
        self.initialize_from(rho, u)

    def init_fields(self) -> None:
        """legacy/lbm_solver.py:1067-1112: rho = 1, u = 0, f = f_eq."""
        self.initialize_fields(1.0, (0.0, 0.0, 0.0))

    def initialize_from(self, rho: Optional[torch.Tensor], u: Optional[torch.Tensor]) -> None:
        """f = f_eq(rho, u) from device tensors ([nz,ny,nx] / [3,nz,ny,nx])."""
        self.ctx.set_current(0)
        self.ctx.init_equilibrium(rho, u)
        self._macro_valid = False
        self._step_count = 0
        self._mass0 = None
        self._mass0 = self.get_diagnostics()["mass"]

    def set_geometry(self, geometry_function: Any) -> None:
        """Accept a Geometry (geometry.py) or a flag-word array [nz,ny,nx]."""
        flags = getattr(geometry_function, "flags", geometry_function)
        flags = np.ascontiguousarray(flags, dtype=np.uint32)
        if self._flags_host is None:
            raise ValueError("solver was built without cell flags; pass the geometry at construction")
        self.ctx.set_flags(flags)
        # the compact layout follows the flags: the state buffers are new
        self._f_views = [DeviceField(t, "population", after_write=self._invalidate, deviation=True, ctx=self.ctx)
                         for t in self.ctx.f]
        self._flags_host = flags
        kind = (flags >> L.POL_KIND_SHIFT) & L.POL_KIND_MASK
        self._solid_t.copy_(torch.from_numpy((kind == L.POL_KIND_SOLID).astype(np.uint8)))
        self._macro_valid = False

    # ---- forces ---------------------------------------------------------------------
    def clear_body_force(self) -> None:
        """legacy/lbm_solver.py:656-660."""
        if self.ctx.force is not None:
            self.ctx.force.zero_()

    def add_force_term(self, force_field: Any) -> None:
        """Add a force-density field ((NX,NY,NZ,3) numpy, DeviceField or [3,nz,ny,nx] tensor)."""
        if self.ctx.force is None:
            raise ValueError("solver built without a force field (SolverConfig.force_field=False)")
        if isinstance(force_field, DeviceField):
            t = force_field.to_torch()
        elif isinstance(force_field, torch.Tensor):
            t = force_field
        else:
            t = torch.from_numpy(np.ascontiguousarray(np.asarray(force_field).transpose(3, 2, 1, 0)))
        self.ctx.force.add_(t.to(self.ctx.device, self.ctx.force.dtype))

    def add_particle_reaction_forces(self, particle_system: Any) -> None:
        """legacy/lbm_solver.py:1478-1483: body_force += the particle system's
        reaction_force_field on the non-solid cells (explicit reaction).  With
        the implicit reaction there is no field to add: the collision applies
        the particles' drag field itself, so this is a no-op."""
        if getattr(particle_system, "reaction", "implicit") != "explicit":
            return
        if self.ctx.force is None:
            raise ValueError("solver built without a force field (SolverConfig.force_field=True is needed)")
        st = self.ctx.lib.pol_add_reaction_forces(self.ctx.handle, particle_system._reaction.data_ptr(),
                                                  L.stream_handle())
        L.check(st, self.ctx.handle, "pol_add_reaction_forces")

    def step_with_two_way_coupling(self, particle_system: Any, dt: float = 1.0, relaxation_factor: float = 0.8,
                                   advance_particles: bool = True) -> None:
        """legacy/lbm_solver.py:1485-1509: clear_body_force (explicit reaction) ->
        compute_two_way_coupling_forces(self.u) -> apply_under_relaxation ->
        add_particle_reaction_forces -> step, then (advance_particles) the
        particle push the reference left commented out.  A coupling already
        made by update_particle_coupling is used as is.  The implicit form
        with the push runs as one native loop (pol_coupled_step), whose
        coupling reads the velocity the previous step stored while colliding
        (the legacy solver's u); the sequence of separate calls reads
        self.u, the moments of the current state."""
        ps = particle_system
        explicit = getattr(ps, "reaction", "implicit") == "explicit"
        if explicit or not advance_particles or self._pending_coupling is ps:
            if self._pending_coupling is not ps:
                if explicit:
                    self.clear_body_force()
                ps.compute_two_way_coupling_forces(self.u)
                ps.apply_under_relaxation(relaxation_factor)
                if explicit:
                    self.add_particle_reaction_forces(ps)
                self._pending_coupling = ps
            self.step()  # consumes the coupling
            if advance_particles:
                ps.advance()
            return
        self._pending_coupling = None
        self._last_particles = ps
        ps.coupled_step(self, relaxation_factor=relaxation_factor)

    def update_particle_coupling(self, particle_system: Any) -> None: ...
    def compute_drag_force(self) -> Any: ...


def validate_lbm_solver(solver: Any) -> bool:
    """isinstance check of src/core/lbm_protocol.py:241-257."""
    return isinstance(solver, LBMSolverProtocol)


class HipMemoryAdapter:
    """Memory adapter carrying the HIP context (attribute contract of cpu_backend.py:207-212)."""

    layout_type = "SoA f[q][z][y][x]"

    def __init__(self, solver: "HipLBMSolver"):
        self.hip_context = _HipContextAdapter(solver)
        self._solver = solver

    @property
    def f(self):
        return self._solver.f

    @property
    def f_new(self):
        return self._solver.f_new

    @property
    def rho(self):
        return self._solver.rho

    @property
    def u(self):
        return self._solver.u

    @property
    def solid(self):
        return self._solver.solid


class _HipContextAdapter:
    """HipContext wrapper that marks solver macros stale after steps."""

    def __init__(self, solver: "HipLBMSolver"):
        self._s = solver
        self.cells = solver.config.cells

    @property
    def config(self) -> SolverConfig:
        return self._s.config

    def step(self, n: int = 1) -> None:
        self._s._advance(n)

    def macroscopic(self) -> None:
        self._s._ensure_macro()

    def synchronize(self) -> None:
        self._s.ctx.synchronize()


class HipLBMSolver:
    """D3Q19 solver on MI355X.

    Parameters come in a SolverConfig; geometry (cell flags) is optional
    (None = all-fluid box with the configured periodicity).
    """

    def __init__(self, config: SolverConfig, flags: Optional[np.ndarray] = None, backend: Optional[ComputeBackend] = None):
        self.config = config
        self.backend = backend if backend is not None else HipBackend()
        if flags is None and not all(config.periodic):
            from .geometry import link_flags

            kind = np.zeros((config.nz, config.ny, config.nx), dtype=np.uint8)
            flags, _ = link_flags(kind, periodic=tuple(bool(p) for p in config.periodic))
        self.ctx = HipContext(config, flags)
        self._flags_host = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint32)
        self.use_les = config.les
        self._macro_valid = False
        self._step_count = 0
        self._mass0: Optional[float] = None
        dev16 = True  # device buffers hold f - w in every precision
        # reference-style field views
        self.rho = DeviceField(self.ctx.rho, "scalar", before_read=self._ensure_macro)
        self.u = DeviceField(self.ctx.u, "vector", before_read=self._ensure_macro)
        self._f_views = [DeviceField(t, "population", after_write=self._invalidate, deviation=dev16, ctx=self.ctx)
                         for t in self.ctx.f]
        self.body_force = DeviceField(self.ctx.force, "vector") if self.ctx.force is not None else None
        self.phase = DeviceField(torch.ones(self.ctx.shape, dtype=config.real_dtype, device=self.ctx.device),
                                 "scalar")
        self._solid_t = torch.zeros(self.ctx.shape, dtype=torch.uint8, device=self.ctx.device)
        if self._flags_host is not None:
            kind = (self._flags_host >> L.POL_KIND_SHIFT) & L.POL_KIND_MASK
            self._solid_t.copy_(torch.from_numpy((kind == L.POL_KIND_SOLID).astype(np.uint8)))
        self.solid = DeviceField(self._solid_t, "scalar")
        self.opposite_dir = OPP.copy()
        self.memory_adapter = HipMemoryAdapter(self)
        self.boundary_manager = None  # boundaries are applied inside the fused step
        self.thermal_coupling_enabled = False
        self._pending_coupling = None  # particle system coupled by update_particle_coupling, not yet stepped
        self._last_particles = None
        # the reference's finite-difference LES (legacy/lbm_solver.py:93-111
        # hands the solver's phase field to the model)
        self.les_model = None
        if config.les_fd:
            from .les import HipLESTurbulenceModel

            self.les_model = HipLESTurbulenceModel(self.ctx)
            self.les_model.set_phase_field(self.phase)
            self.les_nu_sgs = self.les_model.nu_sgs

    # ---- fields ------------------------------------------------------------------
    @property
    def f(self) -> DeviceField:
        """Current populations (post-collision state of the last step)."""
        return self._f_views[self.ctx.current]

    @property
    def f_new(self) -> DeviceField:
        return self._f_views[1 - self.ctx.current]

    @property
    def ux(self) -> DeviceField:
        self._ensure_macro()
        return DeviceField(self.ctx.u[0], "scalar")

    @property
    def uy(self) -> DeviceField:
        self._ensure_macro()
        return DeviceField(self.ctx.u[1], "scalar")

    @property
    def uz(self) -> DeviceField:
        self._ensure_macro()
        return DeviceField(self.ctx.u[2], "scalar")

    @property
    def les_mask(self) -> DeviceField:
        """1 where LES is allowed (filter_paper.py:199-204 turns it off in the paper)."""
        if self._flags_host is None:
            m = np.ones(self.ctx.shape, dtype=np.int32)
        else:
            m = ((self._flags_host & L.POL_FLAG_NO_LES) == 0).astype(np.int32)
        return DeviceField(torch.from_numpy(m).to(self.ctx.device), "scalar")

    @property
    def step_count(self) -> int:
        return self._step_count

    def _invalidate(self) -> None:
        self._macro_valid = False

    def _ensure_macro(self) -> None:
        if not self._macro_valid:
            self.ctx.macroscopic(macro=True)
            self._macro_valid = True

    def _advance(self, n: int) -> None:
        self.ctx.step(n)
        self._step_count += n
        self._macro_valid = False

    # ---- core operations ----------------------------------------------------------
    def step(self) -> None:
        """One fused stream-collide step through the backend (lbm_unified.py:212-222).
        A coupling left by update_particle_coupling acts on this step only."""
        self.backend.execute_collision_streaming(self.memory_adapter, {"tau": self.config.tau, "sync": False})
        self._consume_coupling()

    def run(self, nsteps: int) -> None:
        """nsteps fused steps in one C call (no per-step Python work)."""
        if nsteps <= 0:
            return
        if self._pending_coupling is not None:
            self.step()
            nsteps -= 1
            if nsteps == 0:
                return
        self.backend.execute_collision_streaming(self.memory_adapter,
                                                 {"tau": self.config.tau, "steps": int(nsteps), "sync": False})

    def _consume_coupling(self) -> None:
        ps = self._pending_coupling
        if ps is None:
            return
        self._pending_coupling = None
        self._last_particles = ps
        if getattr(ps, "reaction", "implicit") != "explicit":
            ps.clear_reaction_forces()  # the drag-field footprint of this step's coupling

    def collision_step(self) -> None:
        """Collision is fused with streaming: this runs one full step."""
        self.step()

    def streaming_step(self) -> None:
        """Streaming is fused into collision_step/step; nothing separate to run."""

    def compute_macroscopic_quantities(self) -> None:
        self._macro_valid = False
        self._ensure_macro()

    def apply_boundary_conditions(self) -> None:
        """Walls (halfway bounce-back), inlet and outlet act inside the fused step."""
        self.backend.apply_boundary_conditions(self.memory_adapter)

    def swap_fields(self) -> None:
        """Buffers swap by pointer inside pol_step (legacy/lbm_solver.py:630-654 copied them)."""

    # ---- initialisation -----------------------------------------------------------
    def initialize_fields(self, initial_density: float = 1.0,
                          initial_velocity: Tuple[float, float, float] = (0.0, 0.0, 0.0)) -> None:
        rd = self.config.real_dtype
        rho = torch.full(self.ctx.shape, float(initial_density), dtype=rd, device=self.ctx.device)
        u = torch.empty((3,) + self.ctx.shape, dtype=rd, device=self.ctx.device)
        for i in range(3):
            u[i].fill_(float(initial_velocity[i]))
        self.initialize_from(rho, u)

    def init_fields(self) -> None:
        """legacy/lbm_solver.py:1067-1112: rho = 1, u = 0, f = f_eq."""
        self.initialize_fields(1.0, (0.0, 0.0, 0.0))

    def initialize_from(self, rho: Optional[torch.Tensor], u: Optional[torch.Tensor]) -> None:
        """f = f_eq(rho, u) from device tensors ([nz,ny,nx] / [3,nz,ny,nx])."""
        self.ctx.set_current(0)
        self.ctx.init_equilibrium(rho, u)
        self._macro_valid = False
        self._step_count = 0
        self._mass0 = None
        self._mass0 = self.get_diagnostics()["mass"]

    def set_geometry(self, geometry_function: Any) -> None:
        """Accept a Geometry (geometry.py) or a flag-word array [nz,ny,nx]."""
        flags = getattr(geometry_function, "flags", geometry_function)
        flags = np.ascontiguousarray(flags, dtype=np.uint32)
        if self._flags_host is None:
            raise ValueError("solver was built without cell flags; pass the geometry at construction")
        self.ctx.set_flags(flags)
        # the compact layout follows the flags: the state buffers are new
        self._f_views = [DeviceField(t, "population", after_write=self._invalidate, deviation=True, ctx=self.ctx)
                         for t in self.ctx.f]
        self._flags_host = flags
        kind = (flags >> L.POL_KIND_SHIFT) & L.POL_KIND_MASK
        self._solid_t.copy_(torch.from_numpy((kind == L.POL_KIND_SOLID).astype(np.uint8)))
        self._macro_valid = False

    # ---- forces ---------------------------------------------------------------------
    def clear_body_force(self) -> None:
        """legacy/lbm_solver.py:656-660."""
        if self.ctx.force is not None:
            self.ctx.force.zero_()

    def add_force_term(self, force_field: Any) -> None:
        """Add a force-density field ((NX,NY,NZ,3) numpy, DeviceField or [3,nz,ny,nx] tensor)."""
        if self.ctx.force is None:
            raise ValueError("solver built without a force field (SolverConfig.force_field=False)")
        if isinstance(force_field, DeviceField):
            t = force_field.to_torch()
        elif isinstance(force_field, torch.Tensor):
            t = force_field
        else:
            t = torch.from_numpy(np.ascontiguousarray(np.asarray(force_field).transpose(3, 2, 1, 0)))
        self.ctx.force.add_(t.to(self.ctx.device, self.ctx.force.dtype))

    def add_particle_reaction_forces(self, particle_system: Any) -> None:
        """legacy/lbm_solver.py:1478-1483: body_force += the particle system's
        reaction_force_field on the non-solid cells (explicit reaction).  With
        the implicit reaction there is no field to add: the collision applies
        the particles' drag field itself, so this is a no-op."""
        if getattr(particle_system, "reaction", "implicit") != "explicit":
            return
        if self.ctx.force is None:
            raise ValueError("solver built without a force field (SolverConfig.force_field=True is needed)")
        st = self.ctx.lib.pol_add_reaction_forces(self.ctx.handle, particle_system._reaction.data_ptr(),
                                                  L.stream_handle())
        L.check(st, self.ctx.handle, "pol_add_reaction_forces")

    def step_with_two_way_coupling(self, particle_system: Any, dt: float = 1.0, relaxation_factor: float = 0.8,
                                   advance_particles: bool = True) -> None:
        """legacy/lbm_solver.py:1485-1509.

        Explicit reaction: the reference's sequence clear_body_force ->
        compute_two_way_coupling_forces(u) -> apply_under_relaxation ->
        add_particle_reaction_forces -> step, then (advance_particles) the
        particle push the reference left commented out.  Implicit reaction:
        the same order runs as one native loop (pol_coupled_step), or, when
        update_particle_coupling already coupled this step, as step + push."""
        ps = particle_system
        if getattr(ps, "reaction", "implicit") == "explicit":
            self.clear_body_force()
            ps.compute_two_way_coupling_forces(self.u)
            ps.apply_under_relaxation(relaxation_factor)
            self.add_particle_reaction_forces(ps)
            self._advance(1)
            self._pending_coupling = None
            if advance_particles:
                ps.advance()
            return
        if self._pending_coupling is ps:
            # update_particle_coupling already spread this step's drag field
            self._pending_coupling = None
            self._advance(1)
            if advance_particles:
                ps.advance()
            else:
                ps.clear_reaction_forces()
            return
        if not advance_particles:
            ps.compute_two_way_coupling_forces(self.u)
            ps.apply_under_relaxation(relaxation_factor)
            self._advance(1)
            ps.clear_reaction_forces()
            return
        self._pending_coupling = None
        ps.coupled_step(self, relaxation_factor=relaxation_factor)

    def update_particle_coupling(self, particle_system: Any) -> None:
        """ParticleCoupledLBMProtocol.update_particle_coupling (lbm_protocol.py:232):
        this step's coupling -- compute_two_way_coupling_forces(u),
        apply_under_relaxation(alpha of the particle system) and, for the
        explicit reaction, add_particle_reaction_forces.  The next step()
        consumes it: the implicit drag-field footprint is cleared after that
        step, so it acts on exactly one step; a repeated call recomputes."""
        ps = particle_system
        ps.compute_two_way_coupling_forces(self.u)
        ps.apply_under_relaxation(ps.desc.relax_alpha)
        if getattr(ps, "reaction", "implicit") == "explicit":
            self.add_particle_reaction_forces(ps)
        self._pending_coupling = ps

    def compute_drag_force(self):
        """ParticleCoupledLBMProtocol.compute_drag_force (lbm_protocol.py:236):
        the per-cell force the particles exert on the fluid, as a vector
        DeviceField ((NX, NY, NZ, 3) on the host): the explicit reaction
        field, or for the implicit reaction BV - B u from the drag field and
        the current velocity (zero when no coupling was computed)."""
        ps = self._pending_coupling or self._last_particles
        if ps is None:
            return DeviceField(torch.zeros((3,) + self.ctx.shape, dtype=self.config.real_dtype,
                                           device=self.ctx.device), "vector")
        if getattr(ps, "reaction", "implicit") == "explicit":
            return ps.reaction_force_field
        self._ensure_macro()
        return DeviceField(ps.implicit_reaction(self.ctx.u), "vector")

    # ---- diagnostics -------------------------------------------------------------------
    def get_diagnostics(self) -> dict:
        d = self.ctx.reduce()
        d["step"] = self._step_count
        d["mass_error"] = abs(d["mass"] - self._mass0) / self._mass0 if self._mass0 else 0.0
        return d

    def get_statistics(self) -> dict:
        """visualizer.py:120-193 keys (max/min/avg velocity, total mass)."""
        d = self.get_diagnostics()
        cells = max(d["cells"], 1.0)
        return {"max_velocity": d["max_velocity"], "avg_density": d["mass"] / cells,
                "total_mass": d["mass"], "kinetic_energy": d["kinetic_energy"], "nonfinite": d["nonfinite_cells"]}

    def check_stability(self, max_velocity: float = 0.3) -> bool:
        d = self.get_diagnostics()
        return d["nonfinite_cells"] == 0 and d["max_velocity"] < max_velocity and d["rho_min"] > 0.0

    def get_kinetic_energy(self) -> float:
        return self.get_diagnostics()["kinetic_energy"]

    def get_mass_conservation_error(self) -> float:
        return self.get_diagnostics()["mass_error"]

    def get_memory_usage(self) -> dict:
        out = {"f": sum(t.numel() * t.element_size() for t in self.ctx.f),
               "rho": self.ctx.rho.numel() * self.ctx.rho.element_size(),
               "u": self.ctx.u.numel() * self.ctx.u.element_size()}
        if self.ctx.flags is not None:
            out["cell_flags"] = self.ctx.flags.numel() * 4
        if self.ctx.force is not None:
            out["body_force"] = self.ctx.force.numel() * self.ctx.force.element_size()
        if self.ctx.perm is not None:
            out["permeability"] = self.ctx.perm.numel() * self.ctx.perm.element_size()
        out["total"] = sum(out.values())
        return out

    def optimize_memory_layout(self) -> None:
        """Layout is fixed SoA f[q][z][y][x] (x fastest, 128-byte rows)."""

    def enable_les_turbulence(self, smagorinsky_constant: float = 0.17) -> None:
        """lbm_protocol.py:168: switch the Smagorinsky eddy viscosity on with
        C_s = smagorinsky_constant for the steps enqueued afterwards (the fused
        kernel reads the LES bit per launch; cells flagged LES-off, the filter
        paper, stay laminar)."""
        self.ctx.set_params(les=True, cs_smag=float(smagorinsky_constant))
        self.config = self.ctx.config
        self.use_les = True

    def disable_les_turbulence(self) -> None:
        """Switch the eddy viscosity off again (molecular tau only)."""
        self.ctx.set_params(les=False)
        self.config = self.ctx.config
        self.use_les = False

    def get_velocity_field_for_thermal_coupling(self) -> torch.Tensor:
        """Device velocity [3,nz,ny,nx] (legacy/lbm_solver.py:1195-1206 returned the u field)."""
        self._ensure_macro()
        return self.ctx.u

    def enable_thermal_coupling_output(self, enable: bool = True) -> None:
        self.thermal_coupling_enabled = bool(enable)

    def export_vtk(self, filename: str) -> None:
        """Legacy-VTK structured points with density and velocity (ASCII header, binary big-endian data)."""
        self._ensure_macro()
        rho = self.ctx.rho.float().cpu().numpy()
        u = self.ctx.u.float().cpu().numpy()
        nz, ny, nx = rho.shape
        with open(filename, "wb") as fh:
            fh.write(b"# vtk DataFile Version 3.0\nHIP D3Q19\nBINARY\nDATASET STRUCTURED_POINTS\n")
            fh.write(("DIMENSIONS %d %d %d\nORIGIN 0 0 0\nSPACING 1 1 1\nPOINT_DATA %d\n"
                      % (nx, ny, nz, nx * ny * nz)).encode())
            fh.write(b"SCALARS density float 1\nLOOKUP_TABLE default\n")
            fh.write(rho.astype(">f4").tobytes())
            fh.write(b"\nVECTORS velocity float\n")
            fh.write(np.ascontiguousarray(u.transpose(1, 2, 3, 0)).astype(">f4").tobytes())


class UnifiedLBMSolver(HipLBMSolver):
    """src/core/lbm_unified.py:32-312 shape: backend chosen by name ('hip' or 'auto')."""

    def __init__(self, preferred_backend: Optional[str] = None, config: Optional[SolverConfig] = None,
                 flags: Optional[np.ndarray] = None):
        name = "hip" if preferred_backend in (None, "auto", "hip") else preferred_backend
        backend = create_backend(name)
        super().__init__(config or SolverConfig(224, 224, 224), flags=flags, backend=backend)
        self.tau = self.config.tau

    def get_solver_info(self) -> Dict[str, Any]:
        return {"name": "Unified LBM Solver (HIP)", "backend": self.backend.get_backend_info(),
                "grid_size": (self.config.nx, self.config.ny, self.config.nz),
                "memory_usage_gb": self.backend.estimate_memory_usage(self.config.nx, self.config.ny, self.config.nz,
                                                                      self.config.bytes_per_population),
                "physics_params": {"tau": self.config.tau}}
