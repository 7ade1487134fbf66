"""Cahn-Hilliard phase field with CSF surface tension on the HIP engine.

Mirrors (reference): src/core/multiphase_3d.py:13-579 MultiphaseFlow3D --
same constructor argument (the LBM solver), constants, field names and
method names, with the methods that the class resolves to at run time (the
later definitions of compute_surface_tension_force, :365-381, and step,
:389-414).  The work runs in csrc/multiphase.hip through pol_multiphase_*:
  accumulate_surface_tension_pre_collision  one pass (gradients, curvature,
                                             CSF force, body_force += F/rho)
  step(step_count, precollision_applied)    that pass when the reference
                                             applies the force there, then one
                                             pass for advection, M lap(mu),
                                             phase separation, the copy to phi
                                             and update_density_from_phase
The intermediate fields (grad_phi, grad_mu, normal, curvature,
surface_force) are formed when read rather than stored every step; they
hold what the reference's fields hold after a step.  There is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L
from .fields import DeviceField

# the reference's config values (config/core.py, config/physics.py)
SCALE_LENGTH = 0.14 / 224
SCALE_TIME = (SCALE_LENGTH / 0.01) * 1.2
SURFACE_TENSION_LU = (1.0 * (0.02 * SCALE_TIME / SCALE_LENGTH) ** 2 * SCALE_LENGTH) / 1.0
RHO_WATER = 1.0
RHO_AIR = 1.204 / 965.3
DT = 1.0


class MultiphaseFlow3D:
    """multiphase_3d.py:13 on the device.  `lbm_solver` is a HipLBMSolver
    (its context's rho, u, phase and body force are the fields the phase
    field reads and writes; the body force needs SolverConfig(force_field=True))."""

    def __init__(self, lbm_solver, surface_tension: float = SURFACE_TENSION_LU, rho_water: float = RHO_WATER,
                 rho_air: float = RHO_AIR, dt: float = DT):
        if not torch.cuda.is_available():
            raise RuntimeError("MultiphaseFlow3D needs a HIP device (there is no CPU fallback)")
        self.lbm = lbm_solver
        self.ctx = lbm_solver.ctx if hasattr(lbm_solver, "ctx") else lbm_solver
        self.lib = self.ctx.lib
        cfg = self.ctx.config
        self.nx, self.ny, self.nz = cfg.nx, cfg.ny, cfg.nz
        dev = self.ctx.device
        shape = (self.nz, self.ny, self.nx)
        self.INTERFACE_WIDTH = 2.0
        self.MOBILITY = 0.001
        self.SURFACE_TENSION_COEFF = surface_tension
        self.CAHN_NUMBER = 0.005
        self.BETA = 12.0 * self.SURFACE_TENSION_COEFF / self.INTERFACE_WIDTH
        self.KAPPA = 1.5 * self.SURFACE_TENSION_COEFF * self.INTERFACE_WIDTH
        self.RHO_WATER, self.RHO_AIR, self.DT = rho_water, rho_air, dt
        f32 = torch.float32
        self._phi = [torch.zeros(shape, dtype=f32, device=dev) for _ in range(2)]
        self._phi_new = torch.zeros(shape, dtype=f32, device=dev)
        self._mu = torch.zeros(shape, dtype=f32, device=dev)
        self._lap = torch.zeros(shape, dtype=f32, device=dev)
        self._grad_phi = torch.zeros((3,) + shape, dtype=f32, device=dev)
        self._grad_mu = torch.zeros((3,) + shape, dtype=f32, device=dev)
        self._normal = torch.zeros((3,) + shape, dtype=f32, device=dev)
        self._curv = torch.zeros(shape, dtype=f32, device=dev)
        self._sforce = torch.zeros((3,) + shape, dtype=f32, device=dev)
        phase = getattr(lbm_solver, "phase", None)
        self._phase = phase.to_torch() if isinstance(phase, DeviceField) else \
            torch.zeros(shape, dtype=cfg.real_dtype, device=dev)
        d = L.PolMultiphaseDesc()
        d.sigma, d.mobility, d.interface_width = self.SURFACE_TENSION_COEFF, self.MOBILITY, self.INTERFACE_WIDTH
        d.rho_water, d.rho_air, d.dt = rho_water, rho_air, dt
        b = L.PolMultiphaseBuffers()
        b.phi[0], b.phi[1] = self._phi[0].data_ptr(), self._phi[1].data_ptr()
        b.phi_new = self._phi_new.data_ptr()
        b.mu = self._mu.data_ptr()
        b.phase = self._phase.data_ptr()
        L.check(self.lib.pol_multiphase_bind(self.ctx.handle, C.byref(d), C.byref(b)), self.ctx.handle,
                "pol_multiphase_bind")
        self._cur = 0
        self._steps = 0

    # ---- fields (reference index order [x, y, z]) ------------------------------------
    def _interface_fields(self) -> None:
        """Form grad_phi, grad_mu, normal, curvature and surface_force of the
        current phi and mu (on every read: phi may have been changed in place,
        e.g. by the pouring phase change)."""
        out = L.PolInterfaceFields()
        out.grad_phi, out.grad_mu = self._grad_phi.data_ptr(), self._grad_mu.data_ptr()
        out.normal, out.curvature = self._normal.data_ptr(), self._curv.data_ptr()
        out.surface_force = self._sforce.data_ptr()
        L.check(self.lib.pol_multiphase_interface(self.ctx.handle, 0, C.byref(out), L.stream_handle()),
                self.ctx.handle, "pol_multiphase_interface")

    @property
    def phi(self) -> DeviceField:
        return DeviceField(self._phi[self._cur], "scalar")

    @property
    def phi_new(self) -> DeviceField:
        """The reference's phi_new.  The step keeps only its face values (the
        values it copies into phi); after a step its interior reads as the
        current phi, which is what the reference's holds until phi is
        changed from outside."""
        return DeviceField(self._phi_new, "scalar", before_read=self._sync_phi_new)

    def _sync_phi_new(self) -> None:
        if self._steps:
            c = self._phi[self._cur]
            self._phi_new[1:-1, 1:-1, 1:-1].copy_(c[1:-1, 1:-1, 1:-1])

    @property
    def mu(self) -> DeviceField:
        return DeviceField(self._mu, "scalar")

    @property
    def laplacian_phi(self) -> DeviceField:
        return DeviceField(self._lap, "scalar")

    @property
    def grad_phi(self) -> DeviceField:
        return DeviceField(self._grad_phi, "vector", before_read=self._interface_fields)

    @property
    def grad_mu(self) -> DeviceField:
        return DeviceField(self._grad_mu, "vector", before_read=self._interface_fields)

    @property
    def normal(self) -> DeviceField:
        return DeviceField(self._normal, "vector", before_read=self._interface_fields)

    @property
    def curvature(self) -> DeviceField:
        return DeviceField(self._curv, "scalar", before_read=self._interface_fields)

    @property
    def surface_force(self) -> DeviceField:
        return DeviceField(self._sforce, "vector", before_read=self._interface_fields)

    # ---- initialisation (multiphase_3d.py:56-76, 531-579) ----------------------------------
    def _solid(self) -> torch.Tensor:
        flags = getattr(self.ctx, "flags", None)
        if flags is None:
            return torch.zeros(self._phi[0].shape, dtype=torch.bool, device=self._phi[0].device)
        return ((flags.view(torch.int32) >> L.POL_KIND_SHIFT) & L.POL_KIND_MASK) == L.POL_KIND_SOLID

    def init_phase_field(self) -> None:
        """Every cell air (phi = -1)."""
        self._phi[self._cur].fill_(-1.0)

    def _set_dry_initial_state(self) -> None:
        fluid = ~self._solid()
        self._phi[self._cur].masked_fill_(fluid, -1.0)
        self._phi_new.masked_fill_(fluid, -1.0)

    def update_density_from_phase(self) -> None:
        L.check(self.lib.pol_multiphase_density(self.ctx.handle, L.stream_handle()), self.ctx.handle,
                "pol_multiphase_density")

    def compute_chemical_potential(self) -> None:
        L.check(self.lib.pol_multiphase_chemical_potential(self.ctx.handle, self._lap.data_ptr(), L.stream_handle()),
                self.ctx.handle, "pol_multiphase_chemical_potential")

    def compute_gradients(self) -> None:
        self._interface_fields()

    compute_curvature = compute_gradients
    compute_surface_tension_force = compute_gradients

    def standardize_initial_state(self, force_dry_state: bool = True) -> None:
        if force_dry_state:
            self._set_dry_initial_state()
        self.update_density_from_phase()
        self.compute_chemical_potential()
        self.compute_gradients()
        self.validate_initial_phase_consistency()

    def validate_initial_phase_consistency(self) -> dict:
        """multiphase_3d.py:420-529: range, density consistency, solid count, air ratio."""
        phi = self._phi[self._cur]
        bad = int(((phi < -1.1) | (phi > 1.1)).sum().item())
        if bad:
            raise ValueError("%d phase-field values outside [-1, 1]" % bad)
        fluid = ~self._solid()
        rho = self.ctx.rho.float()
        expected = (self.RHO_WATER * (1.0 + phi) + self.RHO_AIR * (1.0 - phi)) * 0.5
        incons = int((((rho - expected).abs() / expected > 0.05) & fluid).sum().item())
        nfl = int(fluid.sum().item())
        air = float(((phi < -0.5) & fluid).sum().item()) / max(nfl, 1)
        return {"out_of_range": bad, "density_inconsistent": incons, "solid_nodes": int(phi.numel() - nfl),
                "air_ratio": air}

    # ---- the step ---------------------------------------------------------------------
    def _ensure_velocity(self) -> None:
        ens = getattr(self.lbm, "_ensure_macro", None)
        if ens is not None:
            ens()

    def apply_surface_tension(self) -> None:
        """body_force += F / rho on fluid cells (the force of the current phi)."""
        L.check(self.lib.pol_multiphase_interface(self.ctx.handle, L.POL_MP_APPLY, None, L.stream_handle()),
                self.ctx.handle, "pol_multiphase_interface")

    def accumulate_surface_tension_pre_collision(self) -> None:
        """multiphase_3d.py:416-421 in one pass."""
        self.apply_surface_tension()

    def step(self, step_count: int = 0, precollision_applied: bool = False) -> None:
        """multiphase_3d.py:389-414."""
        if (not precollision_applied) and step_count > 10:
            self.apply_surface_tension()
        self.update_phase_field_cahn_hilliard()

    # The reference's step runs these four in a row (multiphase_3d.py:405-414);
    # on the device they are one pass, run by the first of them.
    def update_phase_field_cahn_hilliard(self) -> None:
        """multiphase_3d.py:151-197 together with apply_phase_separation,
        copy_phase_field and update_density_from_phase (one pass)."""
        self._ensure_velocity()
        L.check(self.lib.pol_multiphase_step(self.ctx.handle, L.stream_handle()), self.ctx.handle,
                "pol_multiphase_step")
        self._cur ^= 1
        self._steps += 1

    def apply_phase_separation(self) -> None:
        """Applied inside update_phase_field_cahn_hilliard's pass (multiphase_3d.py:383-401)."""

    def copy_phase_field(self) -> None:
        """The phase field ping-pongs: the pass already made phi current (multiphase_3d.py:418-422)."""

    def get_interface_statistics(self) -> dict:
        """multiphase_3d.py:301-322 (float64 reductions on the device)."""
        self._interface_fields()
        phi = self._phi[self._cur]
        m = phi.abs() < 0.9
        gn = torch.linalg.vector_norm(self._grad_phi.double(), dim=0)
        thick = float((1.0 / (gn + 1e-10))[m].mean().item()) if bool(m.any()) else float("nan")
        return {"interface_volume": float(m.sum().item()) * SCALE_LENGTH ** 3,
                "water_fraction": float((phi > 0).sum().item()) / phi.numel(),
                "interface_thickness": thick * SCALE_LENGTH,
                "max_curvature": float(self._curv.abs().max().item()),
                "surface_tension_magnitude": float(torch.linalg.vector_norm(self._sforce.double(), dim=0).max().item())}
