"""numpy restatement of the reference's body-force producers.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the reference of comparison for
csrc/forcing.hip -- never by the product.

Follows src/physics/precise_pouring.py:83-193 (pour position, region
intensity, apply_pouring_force, apply_gradual_phase_change),
src/physics/pressure_gradient_drive.py:55-279 (target density, density
drive, pressure gradient force and its accumulation) and
src/physics/filter_paper.py:616-746 (particle blocking, dynamic blockage).
Arrays are in the device order of the engine ([nz, ny, nx] scalars,
[3, nz, ny, nx] vectors) and float32; expressions keep the reference's
operation order (numpy performs no contraction); exp is formed in float64
and rounded to float32, as the device does.  Python-float constants
the reference writes as expressions of Python floats are formed in double
and rounded once.
"""
from __future__ import annotations

import numpy as np

F = np.float32


def exp32(x):
    """float32 exp formed in double and rounded once (the device's form)."""
    return np.exp(np.asarray(x, dtype=np.float64)).astype(F)


# ---- pouring ---------------------------------------------------------------------------
def spiral_position(pour_time, spiral_speed, spiral_radius, cx, cy, diameter, nx, ny):
    """precise_pouring.py:83-100 in float32 (pattern 1)."""
    t = F(pour_time) * F(spiral_speed)
    r = F(spiral_radius) * (F(1.0) + F(0.1) * t)
    x = F(cx) + r * np.cos(t)
    y = F(cy) + r * np.sin(t)
    x = max(F(diameter), min(F(nx - diameter), x))
    y = max(F(diameter), min(F(ny - diameter), y))
    return F(x), F(y)


def pour_intensity(shape, cx, cy, radius, height, stream_height=4.0) -> np.ndarray:
    """precise_pouring.py:104-130 for every cell: [nz, ny, nx] float32."""
    nz, ny, nx = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    dx = i.astype(F) - F(cx)
    dy = j.astype(F) - F(cy)
    d = np.sqrt(dx * dx + dy * dy)
    rad = F(radius)
    inside = (d <= rad) & (k <= height) & (k.astype(F) >= F(height - stream_height))
    r = d / rad
    with np.errstate(over="ignore", under="ignore"):
        inten = exp32(F(-0.5) * (r * r))
        decay = exp32((-(height - k)).astype(F) / F(2.0))
    return np.where(inside, inten * decay, F(0)).astype(F)


def pour_force(body_force, solid, cx, cy, radius, height, velocity, flow_rate, dt, stream_height=4.0):
    """apply_pouring_force (precise_pouring.py:132-162): new body force."""
    tot = pour_intensity(solid.shape, cx, cy, radius, height, stream_height)
    sel = (solid == 0) & (tot > 0)
    accel = np.zeros_like(tot)
    if F(dt) > F(1e-8):
        accel = ((F(velocity) * tot) * F(flow_rate)) / F(dt)
    accel = np.minimum(accel, F(10.0))
    bf = body_force.copy()
    bf[0] = np.where(sel, bf[0] + F(0.0), bf[0])
    bf[1] = np.where(sel, bf[1] + F(0.0), bf[1])
    bf[2] = np.where(sel, bf[2] + (-accel), bf[2])
    return bf


def pour_phase(phi, solid, cx, cy, radius, height, flow_rate, dt, stream_height=4.0):
    """apply_gradual_phase_change (precise_pouring.py:164-193): new phi."""
    tot = pour_intensity(solid.shape, cx, cy, radius, height, stream_height)
    sel = (solid == 0) & (tot > 0)
    rate = (F(1.0) - phi) / F(0.05)
    rate = np.maximum(F(-2.0), np.minimum(F(2.0), rate))
    change = ((rate * tot) * F(dt)) * F(flow_rate)
    return np.where(sel, np.maximum(F(-1.0), np.minimum(F(1.0), phi + change)), phi).astype(F)


# ---- pressure drive --------------------------------------------------------------------
def target_density(shape, high=1.8, low=0.4, high_zone=0.8, low_zone=0.2) -> np.ndarray:
    """initialize_target_density (pressure_gradient_drive.py:55-72)."""
    nz = shape[0]
    zr = np.arange(nz).astype(F) / F(nz)
    nh = (zr - F(high_zone)) / F(1.0 - high_zone)
    nd = zr / F(low_zone)
    t = np.where(zr >= F(high_zone), F(1.0) + nh * F(high - 1.0),
                 np.where(zr <= F(low_zone), F(low) + nd * F(1.0 - low), F(1.0))).astype(F)
    return np.broadcast_to(t[:, None, None], shape).copy()


def density_drive(rho, target, solid, rate=0.025, half=False, max_adj=0.001, lo=0.5, hi=2.0):
    """apply_density_drive (:96-122); half: the mixed form (:206-227)."""
    adj = (target - rho) * F(rate)
    if half:
        adj = adj * F(0.5)
    adj = np.where(np.abs(adj) > F(max_adj), np.where(adj > 0, F(max_adj), F(-max_adj)), adj)
    new = np.maximum(F(lo), np.minimum(F(hi), rho + adj))
    return np.where(solid == 0, new, rho).astype(F)


def pressure_force(rho, solid, pf_prev=None, max_force=0.12, cs2=1.0 / 3.0):
    """compute_pressure_gradient (:124-173): [3, nz, ny, nx] (fluid cells; others keep pf_prev)."""
    g = []
    for ax in (2, 1, 0):  # x, y, z of a [z, y, x] array
        r = np.moveaxis(rho, ax, -1)
        d = np.empty_like(r)
        d[..., 1:-1] = (r[..., 2:] - r[..., :-2]) * F(0.5)
        d[..., 0] = r[..., 1] - r[..., 0]
        d[..., -1] = r[..., -1] - r[..., -2]
        g.append(np.moveaxis(d, -1, ax))
    c = F(cs2)
    p = [gc * c for gc in g]
    ok = rho > F(1e-12)
    safe = np.where(ok, rho, F(1))
    f = [np.where(ok, (-pc) / safe, F(0)) for pc in p]
    m = np.sqrt((f[0] * f[0] + f[1] * f[1]) + f[2] * f[2])
    big = m > F(max_force)
    s = np.where(big, F(max_force) / np.where(big, m, F(1)), F(1))
    f = [np.where(big, fc * s, fc) for fc in f]
    out = np.zeros((3,) + rho.shape, F) if pf_prev is None else pf_prev.copy()
    for a in range(3):
        out[a] = np.where(solid == 0, f[a], out[a])
    return out


def accumulate(body_force, pf, solid, half=False):
    """_accumulate_pressure_force_to_body_force (:186-190) / mixed (:276-279)."""
    add = F(0.5) * pf if half else pf
    return np.where((solid == 0)[None], body_force + add, body_force).astype(F)


# ---- filter paper ----------------------------------------------------------------------
def filter_blockage(zone, acc, blockage):
    """update_dynamic_resistance (filter_paper.py:699-746): (acc, blockage)."""
    z = zone == 1
    nb = F(0.9) * (F(1.0) - exp32(F(-0.1) * acc))
    b = np.where(z, F(0.95) * blockage + F(0.05) * nb, blockage).astype(F)
    a = np.where(z, acc * F(0.999), acc).astype(F)
    return a, b


def _mix(seed, p, draw):
    u = np.uint32
    with np.errstate(over="ignore"):
        h = u(seed) ^ (u(p) * u(0x9E3779B9)) ^ (u(draw) * u(0x85EBCA6B))
        h ^= h >> u(16)
        h = h * u(0x7FEB352D)
        h ^= h >> u(15)
        h = h * u(0x846CA68B)
        h ^= h >> u(16)
    return (h >> u(8)).astype(F) * F(1.0 / 16777216.0)


def block_particles(zone, acc, pos, vel, active, scale, seed):
    """block_particles_at_filter (filter_paper.py:616-697) with the build's
    counter-based kick (the reference draws ti.random): (acc, vel)."""
    nz, ny, nx = zone.shape
    acc = acc.copy()
    vel = vel.copy()
    for p in range(pos.shape[1]):
        if active[p] == 0:
            continue
        g = [int(pos[a, p] / F(scale)) for a in range(3)]
        if not (0 <= g[0] < nx and 0 <= g[1] < ny and 0 <= g[2] < nz):
            continue
        for oz in range(-2, 3):
            cz = g[2] + oz
            if 0 <= cz < nz and zone[cz, g[1], g[0]] == 1:
                if vel[2, p] < 0:
                    vz = -vel[2, p] * F(0.3)
                    vel[0, p] = vel[0, p] + (_mix(seed, p, 0) - F(0.5)) * F(0.01)
                    vel[1, p] = vel[1, p] + (_mix(seed, p, 1) - F(0.5)) * F(0.01)
                    vel[2, p] = vz
                    acc[cz, g[1], g[0]] += F(0.01)
                break
    return acc, vel
