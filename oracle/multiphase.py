"""numpy restatement of the reference's multiphase phase-field producer.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's cpu_baseline leg
as the reference of comparison for csrc/multiphase.hip -- never by the product.

Follows src/core/multiphase_3d.py as the class resolves at run time (the
later of each duplicated method wins: compute_surface_tension_force at
:365-381 and step at :389-414; step never calls apply_boundary_conditions
or compute_chemical_potential).  Arrays are in the reference's order
(scalars [nx, ny, nz], vectors [nx, ny, nz, 3]) and float32; every
expression keeps the reference's operation order, numpy performs no
contraction.  Constants written as Python expressions of Python floats
(kappa = 3 sigma W / 8, RHO_WATER - RHO_AIR) are formed in double and then
rounded to float32, as the reference's kernel compiler does with Python
scalars.
"""
from __future__ import annotations

import numpy as np

F = np.float32
C = (slice(1, -1),) * 3
XP, XM = (slice(2, None), slice(1, -1), slice(1, -1)), (slice(None, -2), slice(1, -1), slice(1, -1))
YP, YM = (slice(1, -1), slice(2, None), slice(1, -1)), (slice(1, -1), slice(None, -2), slice(1, -1))
ZP, ZM = (slice(1, -1), slice(1, -1), slice(2, None)), (slice(1, -1), slice(1, -1), slice(None, -2))

# the reference's config values (config/core.py, config/physics.py)
SCALE_LENGTH = 0.14 / 224
SCALE_TIME = (SCALE_LENGTH / 0.01) * 1.2
SURFACE_TENSION_LU = (1.0 * (0.02 * SCALE_TIME / SCALE_LENGTH) ** 2 * SCALE_LENGTH) / 1.0
RHO_WATER = 1.0
RHO_AIR = 1.204 / 965.3
INTERFACE_WIDTH = 2.0
MOBILITY = 0.001
DT = 1.0


def _norm3(v: np.ndarray) -> np.ndarray:
    """Taichi's Vector.norm: sqrt((x x + y y) + z z)."""
    return np.sqrt((v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2])


def laplacian(p: np.ndarray) -> np.ndarray:
    """Interior 7-point Laplacian in the reference's summation order (:88-93, :341-344)."""
    return (((((p[XP] + p[XM]) + p[YP]) + p[YM]) + p[ZP]) + p[ZM]) - F(6.0) * p[C]


def chemical_potential(s: dict, sigma: float = SURFACE_TENSION_LU, width: float = INTERFACE_WIDTH) -> None:
    """multiphase_3d.py:78-108: laplacian_phi and mu on interior cells."""
    phi = s["phi"]
    lap = laplacian(phi)
    s["laplacian_phi"][C] = lap
    p = phi[C]
    kappa = F(3.0 * sigma * width / 8.0)
    s["mu"][C] = (p * p * p - p) + (-kappa) * lap


def gradients(s: dict) -> None:
    """multiphase_3d.py:110-131: grad_phi, grad_mu, normal on interior cells."""
    phi, mu = s["phi"], s["mu"]
    h = F(0.5)
    g = np.stack([(phi[XP] - phi[XM]) * h, (phi[YP] - phi[YM]) * h, (phi[ZP] - phi[ZM]) * h], axis=-1)
    s["grad_phi"][C] = g
    s["grad_mu"][C] = np.stack([(mu[XP] - mu[XM]) * h, (mu[YP] - mu[YM]) * h, (mu[ZP] - mu[ZM]) * h], axis=-1)
    mag = _norm3(g)
    ok = mag > F(1e-10)
    n = np.zeros_like(g)
    np.divide(g, mag[..., None], out=n, where=ok[..., None])
    s["normal"][C] = n


def curvature(s: dict) -> None:
    """multiphase_3d.py:133-149: div n on interior cells with |n| > 1e-10."""
    n = s["normal"]
    h = F(0.5)
    k = ((n[XP][..., 0] - n[XM][..., 0]) * h + (n[YP][..., 1] - n[YM][..., 1]) * h) + (n[ZP][..., 2] - n[ZM][..., 2]) * h
    s["curvature"][C] = np.where(_norm3(n[C]) > F(1e-10), k, F(0))


def surface_tension_force(s: dict, sigma: float = SURFACE_TENSION_LU) -> None:
    """multiphase_3d.py:365-381 (the definition that wins): sigma kappa |grad phi| n
    on interior cells with |phi| < 0.9 and |grad phi| > 1e-10."""
    p = s["phi"][C]
    gm = _norm3(s["grad_phi"][C])
    fm = (F(sigma) * s["curvature"][C]) * gm
    f = fm[..., None] * s["normal"][C]
    ok = (np.abs(p) < F(0.9)) & (gm > F(1e-10))
    s["surface_force"][C] = np.where(ok[..., None], f, F(0))


def apply_surface_tension(s: dict) -> None:
    """multiphase_3d.py:403-411: body_force += F / rho on fluid cells with rho > 1e-10."""
    rho = s["rho"]
    ok = (s["solid"] == 0) & (rho > F(1e-10))
    safe = np.where(ok, rho, F(1))
    acc = s["surface_force"] / safe[..., None]
    s["body_force"] = np.where(ok[..., None], s["body_force"] + acc, s["body_force"])


def update_phase_field(s: dict, mobility: float = MOBILITY, dt: float = DT) -> None:
    """multiphase_3d.py:151-197: upwind -u.grad phi plus M lap mu, explicit
    Euler into phi_new, clamped to [-1, 1] (interior cells)."""
    phi, mu, u = s["phi"], s["mu"], s["u"]
    p = phi[C]
    uu = u[C]
    dx = np.where(uu[..., 0] > 0, p - phi[XM], phi[XP] - p)
    dy = np.where(uu[..., 1] > 0, p - phi[YM], phi[YP] - p)
    dz = np.where(uu[..., 2] > 0, p - phi[ZM], phi[ZP] - p)
    conv = -((uu[..., 0] * dx + uu[..., 1] * dy) + uu[..., 2] * dz)
    diff = F(mobility) * laplacian(mu)
    v = p + F(dt) * (conv + diff)
    s["phi_new"][C] = np.maximum(F(-1), np.minimum(F(1), v))


def phase_separation(s: dict, dt: float = DT) -> None:
    """multiphase_3d.py:383-401: phi_new += -0.001 (phi (phi^2 - 1) - 0.01 lap phi) DT
    where |phi| < 0.99 (interior cells)."""
    p = s["phi"][C]
    chem = p * (p * p - F(1)) - F(0.01) * laplacian(s["phi"])
    pn = s["phi_new"][C]
    s["phi_new"][C] = np.where(np.abs(p) < F(0.99), pn + (F(-0.001) * chem) * F(dt), pn)


def density_from_phase(s: dict, rho_water: float = RHO_WATER, rho_air: float = RHO_AIR) -> None:
    """multiphase_3d.py:413-428: rho and phase from clamped phi (every cell)."""
    pc = np.maximum(F(-1), np.minimum(F(1), s["phi"]))
    s["rho"] = F(rho_air) + (F(rho_water - rho_air) * (pc + F(1))) / F(2)
    s["phase"] = (pc + F(1)) / F(2)


def accumulate_surface_tension_pre_collision(s: dict, sigma: float = SURFACE_TENSION_LU) -> None:
    """multiphase_3d.py:416-421."""
    gradients(s)
    curvature(s)
    surface_tension_force(s, sigma)
    apply_surface_tension(s)


def step(s: dict, step_count: int = 0, precollision_applied: bool = False, sigma: float = SURFACE_TENSION_LU,
         mobility: float = MOBILITY, dt: float = DT, rho_water: float = RHO_WATER, rho_air: float = RHO_AIR) -> None:
    """multiphase_3d.py:389-414 (in place on the state dict)."""
    gradients(s)
    curvature(s)
    surface_tension_force(s, sigma)
    if (not precollision_applied) and step_count > 10:
        apply_surface_tension(s)
    update_phase_field(s, mobility, dt)
    phase_separation(s, dt)
    s["phi"] = s["phi_new"].copy()
    density_from_phase(s, rho_water, rho_air)


def new_state(phi: np.ndarray, u: np.ndarray, rho: np.ndarray, solid: np.ndarray,
              body_force: np.ndarray | None = None, mu: np.ndarray | None = None,
              phi_new: np.ndarray | None = None) -> dict:
    """A state dict with the reference's zero-initialised auxiliary fields."""
    shape = phi.shape
    z3 = np.zeros(shape + (3,), F)
    return {"phi": phi.astype(F), "phi_new": (phi_new if phi_new is not None else np.zeros(shape, F)).astype(F),
            "mu": (mu if mu is not None else np.zeros(shape, F)).astype(F), "laplacian_phi": np.zeros(shape, F),
            "grad_phi": z3.copy(), "grad_mu": z3.copy(), "normal": z3.copy(), "curvature": np.zeros(shape, F),
            "surface_force": z3.copy(), "u": u.astype(F), "rho": rho.astype(F), "solid": solid.astype(np.uint8),
            "body_force": (body_force if body_force is not None else z3).astype(F).copy(),
            "phase": np.zeros(shape, F)}


def interface_statistics(s: dict, scale_length: float = SCALE_LENGTH) -> dict:
    """multiphase_3d.py:301-322 on the state's fields (float64 sums)."""
    phi = s["phi"]
    m = np.abs(phi) < 0.9
    gn = np.linalg.norm(s["grad_phi"].astype(np.float64), axis=3)
    thick = float(np.mean((1.0 / (gn + 1e-10))[m])) if m.any() else float("nan")
    return {"interface_volume": float(np.sum(m)) * scale_length ** 3,
            "water_fraction": float(np.sum(phi > 0)) / phi.size,
            "interface_thickness": thick * scale_length,
            "max_curvature": float(np.max(np.abs(s["curvature"]))),
            "surface_tension_magnitude": float(np.max(np.linalg.norm(s["surface_force"].astype(np.float64), axis=3)))}
