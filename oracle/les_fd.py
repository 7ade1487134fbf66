"""numpy restatement of the reference's finite-difference Smagorinsky eddy
viscosity, LESTurbulenceModel._compute_sgs_from_vector
(src/physics/les_turbulence.py:318-380).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the reference of comparison,
never by the product.

Per interior cell (1 <= i < n - 1 on every axis) where the mask allows LES:
central differences of the velocity times 0.5, the symmetric strain rate
S, |S| = sqrt(2 (S11^2 + S22^2 + S33^2 + 2 (S12^2 + S13^2 + S23^2))),
nu = 0 if |S| < 1e-3 or |phi| < 0.9, else min((C_s * 1)^2 |S|, 0.1); zero
on the box faces and where the mask is 0.  Every operation is evaluated in
the field's dtype (float32 like the reference's Taichi fields) in the
reference's order.  Arrays use the device layout: u [3, nz, ny, nx].
"""
from __future__ import annotations

import numpy as np


def les_fd_viscosity(u: np.ndarray, mask=None, phase=None, cs: float = 0.17, nu_max: float = 0.1,
                     dtype=np.float32) -> np.ndarray:
    u = np.asarray(u, dtype=dtype)
    _, nz, ny, nx = u.shape
    h = dtype(0.5)
    nu = np.zeros((nz, ny, nx), dtype=dtype)
    c = (slice(1, nz - 1), slice(1, ny - 1), slice(1, nx - 1))

    def d(comp, axis):
        # central difference along x (axis 2), y (1) or z (0) of component comp
        sp = [slice(1, nz - 1), slice(1, ny - 1), slice(1, nx - 1)]
        sm = list(sp)
        sp[axis] = slice(sp[axis].start + 1, sp[axis].stop + 1)
        sm[axis] = slice(sm[axis].start - 1, sm[axis].stop - 1)
        return (u[comp][tuple(sp)] - u[comp][tuple(sm)]) * h

    dudx, dudy, dudz = d(0, 2), d(0, 1), d(0, 0)
    dvdx, dvdy, dvdz = d(1, 2), d(1, 1), d(1, 0)
    dwdx, dwdy, dwdz = d(2, 2), d(2, 1), d(2, 0)
    S11, S22, S33 = dudx, dvdy, dwdz
    S12 = h * (dudy + dvdx)
    S13 = h * (dudz + dwdx)
    S23 = h * (dvdz + dwdy)
    two = dtype(2.0)
    mag = np.sqrt(two * (S11 * S11 + S22 * S22 + S33 * S33 + two * (S12 * S12 + S13 * S13 + S23 * S23)))
    phi = np.ones_like(mag) if phase is None else np.asarray(phase, dtype=dtype)[c]
    cs2 = dtype(float(cs) * float(cs))  # (C_s * 1)^2 formed in double, rounded once
    loc = cs2 * mag
    nmax = dtype(nu_max)
    with np.errstate(invalid="ignore"):
        val = np.where(loc < nmax, loc, nmax)
        keep = ~(mag < dtype(1e-3)) & ~(np.abs(phi) < dtype(0.9))
    val = np.where(keep, val, dtype(0))
    if mask is not None:
        val = np.where(np.asarray(mask)[c] != 0, val, dtype(0))
    nu[c] = val
    return nu


def mask_from_flags(flags: np.ndarray) -> np.ndarray:
    """1 where the cell flags allow LES (bit POL_FLAG_NO_LES clear)."""
    return ((np.asarray(flags, dtype=np.uint32) & np.uint32(0x10000000)) == 0).astype(np.int32)
