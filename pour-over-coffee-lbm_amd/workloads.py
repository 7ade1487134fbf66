"""The five lattice workloads of the benchmark and parity suite.

  taylor_green   64^3 fully periodic box, tau 0.6, 3-D Taylor-Green vortex
                 (u0 = 0.05, rho = 1), 1000 steps, plain BGK
  v60            224^3 V60 brew domain: 30 deg cone, outlet hole radius 8,
                 2-cell porous filter (K = 1e-3, C_F = 0.55), gravity 1e-5
                 (Guo), Smagorinsky C_s = 0.17, inlet jet radius 4 at
                 u_z = -0.03 on the top plane, tau 0.6, 2000 steps
  v60_particles  v60 + 50 000 coffee particles (lognormal d, median 1.5
                 cells, sigma 0.4, density ratio 1.2, lower 40 % of the cone)
  precision      v60_particles with 16-bit population storage vs FP64
  adjoint        128^3 scaled V60, 500 steps, checkpoints every 50,
                 per-cell K log-uniform in [1e-4, 1e-2]

Reference constants: C_s of config/core.py:90 (0.17), nu_sgs clamp
les_turbulence.py:375 (0.1), under-relaxation 0.8 of
legacy/lbm_solver.py:1485.  tau 0.6 is the Taylor-Green value; the
reference default TAU_FLUID = 0.53 (config/core.py:64) is not used.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .engine import SolverConfig
from .geometry import Geometry, V60Params, build_v60

SEED = 20261202  # PCG64 seed of every randomised input


def taylor_green_fields(n: int = 64, u0: float = 0.05, nx: Optional[int] = None, ny: Optional[int] = None,
                        nz: Optional[int] = None):
    """rho [nz,ny,nx] = 1 and u [3,nz,ny,nx] of the 3-D Taylor-Green vortex (node x = i)."""
    nx = nx or n
    ny = ny or n
    nz = nz or n
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    kx, ky, kz = 2 * math.pi / nx, 2 * math.pi / ny, 2 * math.pi / nz
    u = np.zeros((3, nz, ny, nx), dtype=np.float64)
    u[0] = u0 * np.sin(kx * x) * np.cos(ky * y) * np.cos(kz * z)
    u[1] = -u0 * np.cos(kx * x) * np.sin(ky * y) * np.cos(kz * z)
    rho = np.ones((nz, ny, nx), dtype=np.float64)
    return rho, u


def taylor_green_ke(n: int, u0: float, tau: float, t: float) -> float:
    """Linear-decay kinetic energy KE(t) = KE0 exp(-2 nu |k|^2 t), |k|^2 = 3 (2 pi/N)^2, KE0 = N^3 u0^2 / 8."""
    nu = (tau - 0.5) / 3.0
    k = 2 * math.pi / n
    return 0.5 * n**3 * u0**2 / 4.0 * math.exp(-2.0 * nu * 3 * k * k * t)


def taylor_green_config(n: int = 64, precision: str = "f32", tau: float = 0.6) -> SolverConfig:
    return SolverConfig(n, n, n, precision=precision, tau=tau, periodic=(True, True, True))


@dataclass
class V60Workload:
    config: SolverConfig
    geometry: Geometry
    steps: int


def v60_config(n: int = 224, precision: str = "f32", *, les: bool = True, force_field: bool = False,
               perm_field: bool = False, particle_field: bool = False, tau: float = 0.6) -> V60Workload:
    params = V60Params() if n == 224 else V60Params.scaled(n)
    geo = build_v60(params)
    cfg = SolverConfig(n, n, n, precision=precision, tau=tau, periodic=(False, False, False), les=les,
                       cs_smag=0.17, nu_sgs_max=0.1, gravity=(0.0, 0.0, -1e-5), darcy_K=1e-3, forch_CF=0.55,
                       inlet_rho=1.0, inlet_u=(0.0, 0.0, -0.03), outlet_rho=1.0, force_field=force_field,
                       perm_field=perm_field, particle_field=particle_field)
    return V60Workload(cfg, geo, 2000)


@dataclass
class ParticleSeed:
    pos: np.ndarray      # [3, P] float64
    radius: np.ndarray   # [P]
    mass: np.ndarray     # [P]


def seed_particles(geo: Geometry, count: int = 50_000, *, median_d: float = 1.5, sigma: float = 0.4,
                   density_ratio: float = 1.2, lower_frac: float = 0.4, seed: int = SEED) -> ParticleSeed:
    """Uniform positions in the lower `lower_frac` of the cone interior (inside the
    paper, fluid cells only), lognormal diameters, mass = ratio * pi d^3 / 6.
    The set is returned ordered by cell (z, y, x): neighbouring particle slots
    then touch neighbouring cells, so the coupling's gathers and atomic
    scatters share cache lines."""
    p = geo.params
    rng = np.random.Generator(np.random.PCG64(seed))
    zb = p.hole_length + max(1, int(math.ceil(p.filter_thickness)))
    zt = p.hole_length + lower_frac * (geo.cone_top_z - p.hole_length)
    n = p.n
    fluid = (geo.kind == 0) & ~geo.porous
    out = np.empty((3, 0))
    while out.shape[1] < count:
        m = 4 * (count - out.shape[1]) + 1024
        cand = np.stack([rng.uniform(1.0, n - 2.0, m), rng.uniform(1.0, n - 2.0, m), rng.uniform(zb, zt, m)])
        ci = cand.astype(np.int64)
        ok = fluid[ci[2], ci[1], ci[0]]
        # keep the 8 interpolation corners in fluid too
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    ok &= fluid[ci[2] + dz, ci[1] + dy, ci[0] + dx]
        out = np.concatenate([out, cand[:, ok]], axis=1)
    pos = out[:, :count]
    d = median_d * np.exp(sigma * rng.standard_normal(count))
    radius = 0.5 * d
    mass = density_ratio * math.pi * d**3 / 6.0
    ci = pos.astype(np.int64)
    order = np.argsort((ci[2] * n + ci[1]) * n + ci[0], kind="stable")
    return ParticleSeed(np.ascontiguousarray(pos[:, order]), radius[order], mass[order])


def permeability_field(geo: Geometry, lo: float = 1e-4, hi: float = 1e-2, seed: int = SEED) -> np.ndarray:
    """Per-cell K, log-uniform in [lo, hi] on porous cells, hi elsewhere (unused there)."""
    rng = np.random.Generator(np.random.PCG64(seed + 5))
    K = np.full(geo.kind.shape, hi, dtype=np.float64)
    npor = int(geo.porous.sum())
    K[geo.porous] = np.exp(rng.uniform(math.log(lo), math.log(hi), npor))
    return K
