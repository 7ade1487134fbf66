"""LESTurbulenceModel's surface on the HIP engine (src/physics/les_turbulence.py:36-390).

Two eddy-viscosity forms exist in this build:
  - the local non-equilibrium-stress form inside the fused step (default,
    SolverConfig(les=True)): no extra field, tau from the cell's own stress;
  - the reference's finite-difference form (SolverConfig(les=True,
    les_fd=True)): nu_sgs from central differences of the velocity field, the
    reference's masks, thresholds and clamp, formed by a device pass
    (pol_les_viscosity) before every step and read by the collision.
This class presents the second on a context that has it: the nu_sgs field,
update_turbulent_viscosity(u), the LES mask and the phase field of the
interface suppression.
"""
from __future__ import annotations

from typing import Any, Optional

import numpy as np
import torch

from . import _lib as L
from .engine import HipContext
from .fields import DeviceField


class HipLESTurbulenceModel:
    def __init__(self, ctx: HipContext):
        if ctx.nu_sgs is None:
            raise ValueError("the finite-difference LES needs SolverConfig(les=True, les_fd=True)")
        self.ctx = ctx

    # ---- fields (les_turbulence.py:97-103) -------------------------------------
    @property
    def nu_sgs(self) -> DeviceField:
        return DeviceField(self.ctx.nu_sgs, "scalar")

    nu_t = nu_sgs

    @property
    def cs(self) -> float:
        return self.ctx.config.cs_smag

    # ---- update (les_turbulence.py:289-316) ----------------------------------
    def update_turbulent_viscosity(self, u_field: Any = None) -> None:
        """nu_sgs from u_field (a vector DeviceField or [3, nz, ny, nx] device
        tensor of the lattice dtype; None: the context's u buffer)."""
        ptr = None
        if u_field is not None:
            t = u_field.to_torch() if isinstance(u_field, DeviceField) else u_field
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == self.ctx.u.dtype
                    and tuple(t.shape) == tuple(self.ctx.u.shape) and t.is_contiguous()):
                raise ValueError("u_field must be a contiguous [3, nz, ny, nx] device tensor of the lattice dtype")
            ptr = t.data_ptr()
        L.check(self.ctx.lib.pol_les_viscosity(self.ctx.handle, ptr, L.stream_handle()), self.ctx.handle,
                "pol_les_viscosity")

    def compute_sgs_viscosity(self, u: Any, v: Any = None, w: Any = None) -> None:
        """les_turbulence.py:116-186: the eddy viscosity of the velocity field
        (the reference's body reads the vector field; v and w are accepted for
        the signature)."""
        self.update_turbulent_viscosity(u)

    # ---- inputs (les_turbulence.py:383-390) -------------------------------------
    def set_mask(self, mask_field: Any) -> None:
        """1 = LES allowed, 0 = off: written into the cell flags' LES-off bit
        (host array (NX, NY, NZ) in the reference's index order, a DeviceField,
        or a [nz, ny, nx] tensor)."""
        ctx = self.ctx
        if ctx.flags is None:
            raise ValueError("the context has no cell flags to carry a LES mask")
        if isinstance(mask_field, DeviceField):
            m = mask_field.to_torch()
        elif isinstance(mask_field, torch.Tensor):
            m = mask_field
        else:
            m = torch.from_numpy(np.ascontiguousarray(np.asarray(mask_field).transpose(2, 1, 0)))
        off = (m.to(ctx.device) == 0).to(torch.int32) * int(L.POL_FLAG_NO_LES)
        ctx.flags.copy_((ctx.flags & ~int(L.POL_FLAG_NO_LES)) | off)

    def set_phase_field(self, phase_field: Optional[Any]) -> None:
        """Phase field of the |phi| < 0.9 suppression (a scalar DeviceField or
        a [nz, ny, nx] device tensor, kept referenced), or None (phi = 1)."""
        ctx = self.ctx
        if phase_field is None:
            ctx.les_phase = None
        else:
            t = phase_field.to_torch() if isinstance(phase_field, DeviceField) else phase_field
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == ctx.u.dtype
                    and tuple(t.shape) == tuple(ctx.shape) and t.is_contiguous()):
                raise ValueError("the phase field must be a contiguous [nz, ny, nx] device tensor of the lattice dtype")
            ctx.les_phase = t
        ctx._bind_keep_state()
