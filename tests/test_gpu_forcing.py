"""Body-force producers on the device (csrc/forcing.hip through pol_pour_*,
pol_density_drive, pol_pressure_force, pol_filter_*, and the classes of
pourover_amd.forcing) against the numpy restatement of the reference's
pouring, pressure drive and filter dynamics (oracle/forcing.py).

Every field compares bit for bit: the pour intensity and the blockage law
form exp in double and round once on both sides."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import forcing as OF
from pourover_amd import _lib as L
from pourover_amd import forcing as FC
from pourover_amd.engine import SolverConfig
from pourover_amd.geometry import link_flags
from pourover_amd.solver import HipLBMSolver

pytestmark = pytest.mark.gpu

NX, NY, NZ = 22, 20, 24
F = np.float32


def _same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype, (what, a.shape, b.shape, a.dtype, b.dtype)
    iv = np.int64 if a.dtype.itemsize == 8 else np.int32
    bad = np.argwhere(a.view(iv) != b.view(iv))
    if bad.size:
        i = tuple(bad[0])
        raise AssertionError("%s differs at %d places, first %s: %r vs %r" % (what, len(bad), i, a[i], b[i]))


def _kind():
    kind = np.zeros((NZ, NY, NX), np.uint8)
    kind[:, :, :2] = L.POL_KIND_SOLID
    kind[16:19, 8:10, 9:12] = L.POL_KIND_SOLID  # inside the pour region
    return kind


def _solver(precision="f32"):
    cfg = SolverConfig(NX, NY, NZ, precision=precision, tau=0.8, periodic=(False, False, False), force_field=True)
    flags, _ = link_flags(_kind())
    fluid = HipLBMSolver(cfg, flags)
    fluid.initialize_fields()
    return fluid


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _solid():
    return (_kind() == L.POL_KIND_SOLID).astype(np.uint8)


# the last centre lies far outside the grid (beyond the 64-bit integer range
# of the pour's bounding box): no cell is poured, and the box stays defined
@pytest.mark.parametrize("center", [(10.0, 9.0), (1.0, 18.5), (1.0e19, -3.0e19)])
def test_pour_force_and_phase(center):
    print("case: pour force / phase centre %s" % (center,), flush=True)
    fluid = _solver()
    rng = np.random.default_rng(3)
    bf = rng.normal(0, 1e-3, (3, NZ, NY, NX)).astype(F)
    bf[0, 17, 12, 10] = F(-0.0)  # a fluid cell of the region: -0 + 0 = +0
    fluid.ctx.force.copy_(_dev(bf))
    pour = FC.PrecisePouringSystem(fluid)
    assert pour.POUR_HEIGHT == NZ - 6 and pour.POUR_DIAMETER_GRID == 8.0 and pour.POUR_VELOCITY == 0.05
    pour.start_pouring(*center, flow_rate=0.8)
    pour.apply_pouring_force(fluid.body_force, fluid.solid, 0.5)
    torch.cuda.synchronize()
    want = OF.pour_force(bf, _solid(), center[0], center[1], 4.0, pour.POUR_HEIGHT, 0.05, 0.8, 0.5)
    got = fluid.ctx.force.cpu().numpy()
    _same(got, want, "body force")
    assert np.signbit(got[0, 17, 12, 10]) == np.signbit(want[0, 17, 12, 10])
    inside = abs(center[0]) < NX and abs(center[1]) < NY
    poured = int((want[2] != bf[2]).sum())
    assert poured > 40 if inside else poured == 0
    phi = rng.uniform(-1, 1, (NZ, NY, NX)).astype(F)
    t = _dev(phi)
    pour.apply_gradual_phase_change(t, fluid.solid, 0.5)
    want_phi = OF.pour_phase(phi, _solid(), center[0], center[1], 4.0, pour.POUR_HEIGHT, 0.8, 0.5)
    _same(t.cpu().numpy(), want_phi, "phi")
    assert pour.pour_time == F(0.5)


def test_spiral_pour_sequence():
    print("case: spiral pour, 6 calls", flush=True)
    fluid = _solver()
    pour = FC.PrecisePouringSystem(fluid)
    pour.start_pouring(11.0, 10.0, flow_rate=1.0, pattern="spiral")
    bf = np.zeros((3, NZ, NY, NX), F)
    t = F(0)
    for i in range(6):
        # the reference's call forms: dt positional (last) or by keyword
        if i % 2:
            pour.apply_pouring_force(fluid.body_force, fluid.solid, dt=0.4)
        else:
            pour.apply_pouring_force(0.4)
        t = F(t + F(0.4))
        x, y = OF.spiral_position(t, 1.0, 5.0, 11.0, 10.0, 8.0, NX, NY)
        bf = OF.pour_force(bf, _solid(), x, y, 4.0, pour.POUR_HEIGHT, 0.05, 1.0, 0.4)
    _same(fluid.ctx.force.cpu().numpy(), bf, "body force after 6 spiral pours")
    info = pour.get_pouring_info()
    assert info["active"] and info["pattern"] == 1 and abs(info["pour_time"] - 2.4) < 1e-6


@pytest.mark.parametrize("mode", ["force", "mixed", "density"])
def test_pressure_drive(mode):
    print("case: pressure drive %s" % mode, flush=True)
    fluid = _solver()
    rng = np.random.default_rng(5)
    rho = rng.uniform(0.9, 1.1, (NZ, NY, NX)).astype(F)
    rho[4, 4, 4] = F(1e-13)  # rho <= 1e-12: zero force
    bf = rng.normal(0, 1e-3, (3, NZ, NY, NX)).astype(F)
    fluid.ctx.rho.copy_(_dev(rho))
    fluid.ctx.force.copy_(_dev(bf))
    fluid._macro_valid = True
    drive = FC.PressureGradientDrive(fluid)
    _same(drive._target.cpu().numpy(), OF.target_density((NZ, NY, NX)), "target density")
    solid = _solid()
    getattr(drive, "activate_%s_drive" % mode)(True)
    drive.apply(1)
    torch.cuda.synchronize()
    if mode == "density":
        _same(fluid.ctx.rho.cpu().numpy(), OF.density_drive(rho, OF.target_density(rho.shape), solid), "rho")
        _same(fluid.ctx.force.cpu().numpy(), bf, "body force untouched")
        return
    pf = OF.pressure_force(rho, solid)
    _same(drive._pforce.cpu().numpy(), pf, "pressure force")
    _same(fluid.ctx.force.cpu().numpy(), OF.accumulate(bf, pf, solid, half=(mode == "mixed")), "body force")
    st = drive.get_statistics()
    p = rho[solid == 0].astype(np.float64) * (1.0 / 3.0)
    assert st["max_pressure"] == pytest.approx(p.max(), rel=1e-7) and st["min_pressure"] == pytest.approx(p.min(), rel=1e-7)


def test_filter_dynamics():
    print("case: filter blockage / particle blocking", flush=True)
    fluid = _solver()
    zone = np.zeros((NZ, NY, NX), np.uint8)
    zone[5] = 1
    filt = FC.FilterPaperSystem(fluid, zone, seed=11)
    rng = np.random.default_rng(9)
    acc = rng.uniform(0, 20, zone.shape).astype(F)
    blk = rng.uniform(0, 0.5, zone.shape).astype(F)
    filt._acc.copy_(_dev(acc))
    filt._blockage.copy_(_dev(blk))
    filt.update_dynamic_resistance()
    a, b = OF.filter_blockage(zone, acc, blk)
    _same(filt._acc.cpu().numpy(), a, "accumulated particles")
    _same(filt._blockage.cpu().numpy(), b, "blockage")
    # particles at distinct columns: half above the paper moving down
    n = 64
    s = 0.25
    cols = rng.choice(NX * NY, n, replace=False)
    gx, gy = (cols % NX).astype(F), (cols // NX).astype(F)
    gz = rng.integers(3, 9, n).astype(F)
    pos = (np.stack([gx, gy, gz]) + F(0.3)) * F(s)
    vel = rng.normal(0, 0.1, (3, n)).astype(F)
    active = (rng.random(n) < 0.9).astype(np.int32)
    acc0 = filt._acc.cpu().numpy()
    tp, tv, ta = _dev(pos.astype(F)), _dev(vel), _dev(active)
    filt.block_particles_at_filter(tp, tv, None, ta, scale=s)
    seed = (11 * 1000003 + 0) & 0xFFFFFFFF
    wa, wv = OF.block_particles(zone, acc0, pos.astype(F), vel, active, s, seed)
    _same(tv.cpu().numpy(), wv, "particle velocities")
    _same(filt._acc.cpu().numpy(), wa, "accumulated particles after blocking")
    assert (wv[2] != vel[2]).sum() > 5


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("mode", ["force", "mixed"])
def test_pressure_drive_with_surface_tension_in_one_pass(mode, precision):
    """PressureGradientDrive.apply_with_surface_tension (one pass) against the
    pressure drive followed by the separate surface-tension pass: pressure
    force and body force bit for bit, on a droplet interface with random
    density (one cell at rho <= 1e-10) and planted -0.0 components, in FP32
    and FP64 storage.  The separate surface-tension call is disabled before
    the one-pass call, so the test fails if that call falls back to it."""
    from pourover_amd import multiphase as MP

    print("case: pressure drive + surface tension in one pass, %s %s" % (mode, precision), flush=True)
    fluid = _solver(precision)
    RT = np.float64 if precision == "f64" else F
    rng = np.random.default_rng(9)
    rho = rng.uniform(0.9, 1.1, (NZ, NY, NX)).astype(RT)
    rho[12, 10, 11] = RT(1e-11)
    bf = rng.normal(0, 1e-3, (3, NZ, NY, NX)).astype(RT)
    bf[:, 5, 5, 5] = RT(-0.0)
    bf[2, 12, 10, 12] = RT(-0.0)
    z, y, x = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    r = np.sqrt((x - 11.3) ** 2 + (y - 10.1) ** 2 + (z - 11.7) ** 2)
    phi = np.tanh((6.0 - r) / 2.0).astype(F)
    mp = MP.MultiphaseFlow3D(fluid)
    mp._phi[mp._cur].copy_(_dev(phi))
    fluid.ctx.rho.copy_(_dev(rho))
    fluid._macro_valid = True
    drive = FC.PressureGradientDrive(fluid)
    getattr(drive, "activate_%s_drive" % mode)(True)
    fluid.ctx.force.copy_(_dev(bf))
    drive.apply(3)
    mp.accumulate_surface_tension_pre_collision()
    torch.cuda.synchronize()
    pf_two, bf_two = drive._pforce.cpu().numpy(), fluid.ctx.force.cpu().numpy()
    assert np.abs(bf_two - bf).max() > 1e-4            # the interface carries a force
    drive._pforce.zero_()
    fluid.ctx.force.copy_(_dev(bf))

    def _separate_pass_taken():
        raise AssertionError("apply_with_surface_tension fell back to the separate passes")

    mp.accumulate_surface_tension_pre_collision = _separate_pass_taken
    drive.apply_with_surface_tension(mp, 3)
    torch.cuda.synchronize()
    _same(drive._pforce.cpu().numpy(), pf_two, "pressure force")
    _same(fluid.ctx.force.cpu().numpy(), bf_two, "body force")
