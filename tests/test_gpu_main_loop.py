"""The reference's main-loop step (main.py:769-839) composed from the device
parts -- clear body force, pour force and phase change, pressure drive,
pre-collision surface tension, stream-collide writing rho / u, filter
blockage, phase step -- on a 40^3 V60 cup, against the numpy restatements
of the same producers (oracle/forcing.py, oracle/multiphase.py) fed the
device's velocity after each fluid step.

Every part is bit-exact, so the composed step is too: the body force
before each fluid step, phi and rho after 14 steps compare bit for bit."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import forcing as OF
from oracle import multiphase as OM
from pourover_amd import _lib as L
from pourover_amd import forcing as FC
from pourover_amd import multiphase as MP
from pourover_amd import workloads as WL
from pourover_amd.engine import HipContext

pytestmark = pytest.mark.gpu

F = np.float32


def _same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype, (what, a.shape, b.shape, a.dtype, b.dtype)
    bad = np.argwhere(a.view(np.int32) != b.view(np.int32))
    if bad.size:
        i = tuple(bad[0])
        raise AssertionError("%s differs at %d places, first %s: %r vs %r" % (what, len(bad), i, a[i], b[i]))


def _ref(t):  # device [z, y, x] -> reference [x, y, z]
    return np.ascontiguousarray(t.detach().cpu().numpy().transpose(2, 1, 0))


def _refv(t):  # device [3, z, y, x] -> reference [x, y, z, 3]
    return np.ascontiguousarray(t.detach().cpu().numpy().transpose(3, 2, 1, 0))


@pytest.mark.parametrize("fused", [False, True])
def test_main_loop_step_matches_oracles(fused):
    """fused: the pressure drive and the surface tension as one pass
    (PressureGradientDrive.apply_with_surface_tension) from step 11 on."""
    n = 40
    print("case: main-loop step V60 %d^3, 14 steps, fused=%s" % (n, fused), flush=True)
    wl = WL.v60_config(n, "f32", force_field=True)
    ctx = HipContext(wl.config, wl.geometry.flags)
    ctx.init_equilibrium()
    mp = MP.MultiphaseFlow3D(ctx)
    mp.standardize_initial_state(force_dry_state=True)
    pour = FC.PrecisePouringSystem(ctx)
    pour.start_pouring(pattern="spiral")
    drive = FC.PressureGradientDrive(ctx)
    drive.activate_force_drive(True)
    paper = FC.FilterPaperSystem(ctx, wl.geometry.porous)
    solid_dev = (wl.geometry.kind == L.POL_KIND_SOLID).astype(np.uint8)   # [z, y, x]
    s = OM.new_state(_ref(mp._phi[mp._cur]), np.zeros((n, n, n, 3), F), _ref(ctx.rho),
                     solid_dev.transpose(2, 1, 0), mu=_ref(mp._mu), phi_new=_ref(mp._phi_new))
    t = F(0)
    for step in range(1, 15):
        # device
        ctx.force.zero_()
        pour.apply_pouring_force(1.0)
        pour.apply_gradual_phase_change(mp._phi[mp._cur], None, 1.0)
        if fused and step > 10:
            drive.apply_with_surface_tension(mp, step)
        else:
            drive.apply(step)
            if step > 10:
                mp.accumulate_surface_tension_pre_collision()
        bf_dev = ctx.force.cpu().numpy()
        ctx.step(1, L.POL_WANT_MACRO)
        paper.update_dynamic_resistance()
        mp.step(step, precollision_applied=True)
        # oracle twin (device order for the producers, reference order for the phase field)
        t = F(t + F(1.0))
        x, y = OF.spiral_position(t, 1.0, 5.0, n // 2, n // 2, pour.POUR_DIAMETER_GRID, n, n)
        bf = OF.pour_force(np.zeros((3, n, n, n), F), solid_dev, x, y, 4.0, pour.POUR_HEIGHT, 0.05, 1.0, 1.0)
        phi_dev = OF.pour_phase(np.ascontiguousarray(s["phi"].transpose(2, 1, 0)), solid_dev, x, y, 4.0,
                                pour.POUR_HEIGHT, 1.0, 1.0)
        s["phi"] = np.ascontiguousarray(phi_dev.transpose(2, 1, 0))
        rho_dev = np.ascontiguousarray(s["rho"].transpose(2, 1, 0))
        bf = OF.accumulate(bf, OF.pressure_force(rho_dev, solid_dev), solid_dev)
        s["body_force"] = np.ascontiguousarray(bf.transpose(3, 2, 1, 0))
        if step > 10:
            OM.accumulate_surface_tension_pre_collision(s)
        _same(_refv(torch.from_numpy(bf_dev)), s["body_force"], "body force before the fluid step %d" % step)
        s["u"] = _refv(ctx.u)
        OM.step(s, step, precollision_applied=True)
    _same(_ref(mp._phi[mp._cur]), s["phi"], "phi after 14 steps")
    _same(_ref(ctx.rho), s["rho"], "rho after 14 steps")
    wet = s["phi"] > 0
    assert wet.sum() > 0 and np.isfinite(_refv(ctx.u)).all()
    assert abs(pour.pour_time - F(14.0)) == 0
