"""Phase field on the device (csrc/multiphase.hip through pol_multiphase_*,
and pourover_amd.multiphase.MultiphaseFlow3D) against the numpy
restatement of the reference's multiphase_3d.py (oracle/multiphase.py) on
the same seeded fields.

The kernels evaluate the reference's float32 expressions in its order with
contraction off and correctly rounded division and square root, so every
field compares bit for bit: gradients, normals, curvature, CSF force, the
body-force update, chemical potential, phi after steps, rho and phase --
including a run of the reference's main-loop order (clear force, pre-
collision surface tension, fluid step, multiphase step) on the V60-like
solid mask, with the oracle fed the device's velocity."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import multiphase as OM
from pourover_amd import _lib as L
from pourover_amd import multiphase as MP
from pourover_amd.engine import SolverConfig
from pourover_amd.geometry import link_flags
from pourover_amd.solver import HipLBMSolver

pytestmark = pytest.mark.gpu

NX, NY, NZ = 28, 24, 20
F = np.float32


def _xyz(t):
    return np.ascontiguousarray(t.detach().cpu().numpy().transpose(2, 1, 0))


def _vxyz(t):
    return np.ascontiguousarray(t.detach().cpu().numpy().transpose(3, 2, 1, 0))


def _dev(a, like):
    a = np.asarray(a)
    t = a.transpose(3, 2, 1, 0) if a.ndim == 4 else a.transpose(2, 1, 0)
    return torch.from_numpy(np.ascontiguousarray(t)).to(like.device, like.dtype)


def _same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype, (what, a.shape, b.shape, a.dtype, b.dtype)
    it = np.int64 if a.dtype == np.float64 else np.int32
    bad = np.argwhere(a.view(it) != b.view(it))
    if bad.size:
        i = tuple(bad[0])
        raise AssertionError("%s differs at %d places, first %s: %r vs %r" % (what, len(bad), i, a[i], b[i]))


def _kind(nx=NX):
    kind = np.zeros((NZ, NY, nx), np.uint8)  # device order [z, y, x]
    kind[:, :, :3] = L.POL_KIND_SOLID          # a wall slab at low x
    kind[:2, 10:14, 10:14] = L.POL_KIND_SOLID  # a post
    return kind


def _solver(precision="f32", nx=NX):
    cfg = SolverConfig(nx, NY, NZ, precision=precision, tau=0.8, periodic=(False, False, False), force_field=True,
                       gravity=(0.0, 0.0, -2e-5))
    flags, _ = link_flags(_kind(nx))
    fluid = HipLBMSolver(cfg, flags)
    fluid.initialize_fields()
    return fluid


def _phi0(seed, nx=NX):
    rng = np.random.default_rng(seed)
    x, y, z = np.meshgrid(np.arange(nx), np.arange(NY), np.arange(NZ), indexing="ij")
    r = np.sqrt((x - 15.3) ** 2 + (y - 11.7) ** 2 + (z - 9.2) ** 2)
    phi = np.tanh((6.5 - r) / 2.0) + rng.normal(0, 0.01, r.shape)
    return np.clip(phi, -1, 1).astype(F)


def _setup(seed=1, precision="f32", nx=NX):
    fluid = _solver(precision, nx)
    mp = MP.MultiphaseFlow3D(fluid)
    rng = np.random.default_rng(seed + 100)
    phi = _phi0(seed, nx)
    mp._phi[0].copy_(_dev(phi, mp._phi[0]))
    phi_new = rng.uniform(-1, 1, phi.shape).astype(F)
    mp._phi_new.copy_(_dev(phi_new, mp._phi_new))
    mu = (rng.normal(0, 0.05, phi.shape)).astype(F)
    mp._mu.copy_(_dev(mu, mp._mu))
    solid = _xyz(torch.from_numpy(_kind(nx) == L.POL_KIND_SOLID)).astype(np.uint8)
    rho = rng.uniform(0.5, 1.5, phi.shape).astype(F)
    rho[5:8, 3:6, 2:5] = F(1e-12)
    u = (rng.uniform(-0.03, 0.03, phi.shape + (3,))).astype(F)
    u[..., 0][rng.random(phi.shape) < 0.1] = F(0)  # exercise the u == 0 branch of the upwind choice
    bf = (rng.normal(0, 1e-4, phi.shape + (3,))).astype(F)
    bf[::3, ::2, ::5, 1] = F(-0.0)  # -0.0 turns into +0.0 where the reference adds +0 / rho
    bf[1::4, ::3, ::2, 2] = F(0.0)
    ctx = fluid.ctx
    ctx.rho.copy_(_dev(rho, ctx.rho))
    ctx.u.copy_(_dev(u, ctx.u))
    ctx.force.copy_(_dev(bf, ctx.force))
    fluid._macro_valid = True  # the planted rho / u are the state the phase field reads
    s = OM.new_state(phi, u, rho, solid, body_force=bf, mu=mu, phi_new=phi_new)
    torch.cuda.synchronize()
    return fluid, mp, s


def test_interface_fields_match_oracle():
    print("case: multiphase interface fields %dx%dx%d" % (NX, NY, NZ), flush=True)
    fluid, mp, s = _setup(1)
    mp.compute_gradients()
    OM.gradients(s)
    OM.curvature(s)
    OM.surface_tension_force(s)
    _same(_vxyz(mp._grad_phi), s["grad_phi"], "grad_phi")
    _same(_vxyz(mp._grad_mu), s["grad_mu"], "grad_mu")
    _same(_vxyz(mp._normal), s["normal"], "normal")
    _same(_xyz(mp._curv), s["curvature"], "curvature")
    _same(_vxyz(mp._sforce), s["surface_force"], "surface_force")
    assert np.abs(s["surface_force"]).max() > 0 and np.abs(s["curvature"]).max() > 0


def test_chemical_potential_matches_oracle():
    print("case: multiphase chemical potential", flush=True)
    fluid, mp, s = _setup(2)
    mp.compute_chemical_potential()
    OM.chemical_potential(s)
    _same(_xyz(mp._mu), s["mu"], "mu")
    _same(_xyz(mp._lap), s["laplacian_phi"], "laplacian_phi")


def test_apply_surface_tension_matches_oracle():
    print("case: multiphase pre-collision surface tension", flush=True)
    fluid, mp, s = _setup(3)
    mp.accumulate_surface_tension_pre_collision()
    OM.accumulate_surface_tension_pre_collision(s)
    _same(_vxyz(fluid.ctx.force), s["body_force"], "body_force")


@pytest.mark.parametrize("precollision,nx", [(False, NX), (True, NX), (False, 27)])
def test_steps_match_oracle(precollision, nx):
    """nx = 27: an odd row length (rows not aligned to 16 bytes)."""
    print("case: multiphase 3 steps, precollision_applied=%s, nx=%d" % (precollision, nx), flush=True)
    fluid, mp, s = _setup(4, nx=nx)
    for n in range(3):
        sc = 9 + 2 * n  # 9, 11, 13: the force is applied from step_count 11 on
        mp.step(sc, precollision_applied=precollision)
        OM.step(s, sc, precollision_applied=precollision)
    torch.cuda.synchronize()
    _same(_xyz(mp._phi[mp._cur]), s["phi"], "phi")
    _same(_xyz(fluid.ctx.rho), s["rho"], "rho")
    _same(_xyz(mp._phase), s["phase"], "phase")
    _same(_vxyz(fluid.ctx.force), s["body_force"], "body_force")
    _same(mp.phi.to_numpy(), s["phi"], "phi field view")
    # the reference's phi_new after a step: interior = phi, faces untouched
    _same(mp.phi_new.to_numpy(), s["phi_new"], "phi_new")


def test_density_and_statistics():
    print("case: multiphase density / statistics", flush=True)
    fluid, mp, s = _setup(5)
    mp.update_density_from_phase()
    OM.density_from_phase(s)
    _same(_xyz(fluid.ctx.rho), s["rho"], "rho")
    _same(_xyz(mp._phase), s["phase"], "phase")
    st = mp.get_interface_statistics()
    OM.gradients(s)
    OM.curvature(s)
    OM.surface_tension_force(s)
    want = OM.interface_statistics(s)
    for k in want:
        assert st[k] == pytest.approx(want[k], rel=1e-6, abs=1e-30), k


def test_main_loop_order_matches_oracle():
    """main.py:769-839: clear body force, (step > 10) surface tension before
    collision, the fluid step, then multiphase.step(precollision_applied=True)."""
    print("case: multiphase main-loop order, 14 steps", flush=True)
    fluid = _solver()
    mp = MP.MultiphaseFlow3D(fluid)
    mp.init_phase_field()
    phi = _phi0(7)
    mp.phi.from_numpy(phi)
    mp.standardize_initial_state(force_dry_state=False)
    solid = _xyz(torch.from_numpy(_kind() == L.POL_KIND_SOLID)).astype(np.uint8)
    s = OM.new_state(phi, np.zeros(phi.shape + (3,), F), _xyz(fluid.ctx.rho), solid)
    OM.density_from_phase(s)
    OM.chemical_potential(s)
    _same(_xyz(fluid.ctx.rho), s["rho"], "initial rho")
    for step_count in range(1, 15):
        fluid.ctx.force.zero_()
        s["body_force"] = np.zeros_like(s["body_force"])
        if step_count > 10:
            mp.accumulate_surface_tension_pre_collision()
            OM.accumulate_surface_tension_pre_collision(s)
            _same(_vxyz(fluid.ctx.force), s["body_force"], "body force at step %d" % step_count)
        fluid.step()
        s["u"] = fluid.u.to_numpy()
        mp.step(step_count, precollision_applied=True)
        OM.step(s, step_count, precollision_applied=True)
    _same(mp.phi.to_numpy(), s["phi"], "phi after 14 steps")
    _same(_xyz(fluid.ctx.rho), s["rho"], "rho after 14 steps")
    assert np.abs(s["u"]).max() > 0


def test_f64_context():
    print("case: multiphase on an FP64 context", flush=True)
    fluid, mp, s = _setup(6, precision="f64")
    mp.step(12)
    OM.step(s, 12)
    _same(_xyz(mp._phi[mp._cur]), s["phi"], "phi")
    np.testing.assert_array_equal(_xyz(fluid.ctx.rho).astype(F), s["rho"])
    # FP64 body force against the float32 one: the float32 sum rounds at
    # about 6e-8 of the force scale (cancellation leaves only that)
    scale = float(np.abs(s["body_force"]).max())
    np.testing.assert_allclose(_vxyz(fluid.ctx.force), s["body_force"], rtol=1e-6, atol=1.2e-7 * scale)
