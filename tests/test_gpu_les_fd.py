"""The reference's finite-difference eddy viscosity (POL_LES_FD,
les_turbulence.py:318-380) on the device against the oracle.

- pol_les_viscosity against oracle/les_fd.py on the same float32 velocity,
  LES mask (the filter paper's cells) and phase field: bit for bit (both
  evaluate the reference's float32 expressions in its order, without
  contraction, with correctly rounded square roots);
- a 48^3 cup run in this mode (nu_sgs from the velocity the previous step
  wrote, tau = clamp(tau + 3 nu, 0.55, 1.90)) against the C oracle doing the
  same: FP32 within 1e-5 relative L2 on rho and u;
- the LESTurbulenceModel surface (update_turbulent_viscosity, set_mask,
  set_phase_field) on HipLBMSolver.
"""
from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest
import torch

from conftest import rel_l2
from oracle.les_fd import les_fd_viscosity, mask_from_flags
from pourover_amd import _lib as L
from pourover_amd import workloads as WL
from pourover_amd.engine import HipContext
from pourover_amd.solver import HipLBMSolver

pytestmark = pytest.mark.gpu


def _fd_workload(n):
    wl = WL.v60_config(n, "f32")
    return wl, replace(wl.config, les_fd=True)


def test_les_viscosity_bit_exact_against_the_oracle(hip):
    n = 48
    wl, cfg = _fd_workload(n)
    flags = wl.geometry.flags
    print("case: pol_les_viscosity 48^3 vs oracle", flush=True)
    ctx = HipContext(cfg, flags)
    rng = np.random.default_rng(17)
    u = (rng.standard_normal((3, n, n, n)) * 0.01).astype(np.float32)
    u[:, 10:20] *= np.float32(0.01)  # a band below the |S| < 1e-3 cut
    u[:, 30:34] *= np.float32(1000.0)  # a band above the 0.1 clamp
    phase = rng.uniform(-1.0, 1.0, size=(n, n, n)).astype(np.float32)
    ctx.u.copy_(torch.from_numpy(u))
    ctx.les_phase = torch.from_numpy(phase).cuda()
    ctx._bind_keep_state()
    st = ctx.lib.pol_les_viscosity(ctx.handle, None, L.stream_handle())
    assert st == L.POL_OK
    torch.cuda.synchronize()
    got = ctx.nu_sgs.cpu().numpy()
    want = les_fd_viscosity(u, mask=mask_from_flags(flags), phase=phase, cs=cfg.cs_smag, nu_max=cfg.nu_sgs_max)
    assert (want > 0).sum() > 1000 and (want == 0).sum() > 1000 and (want == 0.1).sum() > 0
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_fd_les_run_against_the_oracle(hip, oracle):
    n, steps = 48, 100
    wl, cfg = _fd_workload(n)
    flags = wl.geometry.flags
    print("case: FD-LES 48^3 cup, %d steps" % steps, flush=True)
    ctx = HipContext(cfg, flags)
    ctx.init_equilibrium()
    ctx.macroscopic(macro=True)  # u of the initial state: the first step's nu reads it
    ctx.step(steps)
    ctx.macroscopic(macro=True)
    torch.cuda.synchronize()
    rho_g, u_g = ctx.rho.double().cpu().numpy(), ctx.u.double().cpu().numpy()
    nu_g = ctx.nu_sgs.double().cpu().numpy()

    d = oracle.desc_from_dict(cfg.oracle_dict(True))
    f = oracle.init_equilibrium(n, n, n)
    _, u_prev, _ = oracle.moments(d, f, flags)
    mask = mask_from_flags(flags)
    for _ in range(steps):
        nu = les_fd_viscosity(u_prev, mask=mask, cs=cfg.cs_smag, nu_max=cfg.nu_sgs_max, dtype=np.float64)
        f, _, u_prev, _ = oracle.step(d, f, flags, want_macro=True, nu=nu)
    rho_o, u_o, _ = oracle.moments(d, f, flags)
    fluid = wl.geometry.kind != L.POL_KIND_SOLID
    eu, er = rel_l2(u_g[:, fluid], u_o[:, fluid]), rel_l2(rho_g[fluid], rho_o[fluid])
    print("FD-LES rel L2 u %.3e rho %.3e, cells with nu > 0: %d" % (eu, er, int((nu_g > 0).sum())))
    assert (nu_g > 0).sum() > 0  # the jet shears the fluid above the cut
    assert eu < 1e-5 and er < 1e-5
    # and the mode changes the flow: the local-stress form gives another field
    d_loc = oracle.desc_from_dict(wl.config.oracle_dict(True))
    g = oracle.init_equilibrium(n, n, n)
    for _ in range(steps):
        g, _, _, _ = oracle.step(d_loc, g, flags)
    _, u_loc, _ = oracle.moments(d_loc, g, flags)
    assert rel_l2(u_loc[:, fluid], u_o[:, fluid]) > 1e-4


def test_les_model_surface_on_the_solver(hip):
    n = 40
    wl, cfg = _fd_workload(n)
    print("case: LES model surface, 40^3 cup", flush=True)
    solver = HipLBMSolver(cfg, wl.geometry.flags)
    solver.init_fields()
    solver.compute_macroscopic_quantities()
    solver.run(30)
    les = solver.les_model
    assert les is not None and les.nu_sgs.shape == (n, n, n)
    # a uniform shear through the reference-order field: nu = C_s^2 gamma inside
    z = torch.arange(n, device="cuda", dtype=torch.float32).view(-1, 1, 1)
    u = torch.zeros((3, n, n, n), device="cuda", dtype=torch.float32)
    u[0] = 0.01 * z
    les.set_mask(np.ones((n, n, n), np.int32))  # LES everywhere, the paper included
    les.update_turbulent_viscosity(u)
    nu = les.nu_sgs.to_numpy()
    assert nu.shape == (n, n, n)
    inner = nu[1:-1, 1:-1, 1:-1]
    assert np.allclose(inner, cfg.cs_smag ** 2 * 0.01, rtol=2e-6)
    # the mask switches LES off where it is 0
    m = np.ones((n, n, n), np.int32)
    m[10, 12, 14] = 0
    les.set_mask(m)
    les.update_turbulent_viscosity(u)
    assert les.nu_sgs.to_numpy()[10, 12, 14] == 0.0
    # a phase field with |phi| < 0.9 everywhere suppresses it
    les.set_phase_field(torch.full((n, n, n), 0.5, device="cuda", dtype=torch.float32))
    les.update_turbulent_viscosity(u)
    assert float(solver.ctx.nu_sgs.abs().max()) == 0.0
