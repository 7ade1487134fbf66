"""Size-independent properties of the thermal and phase-field rows at the
full 224^3 grid (the oracles run at small sizes; these laws hold at any
size): a Gaussian temperature blob keeps its heat and spreads its variance
at 2 D t with the lattice diffusivity D = (tau - 1/2) / 4 of the D3Q7
weights; a planar water / air interface at rest stays exactly planar under
the phase step (outside the reach of the frozen face cells) and carries no
CSF force away from the faces."""
from __future__ import annotations

import pytest
import torch

from pourover_amd import multiphase as MP
from pourover_amd import thermal as TH
from pourover_amd.engine import SolverConfig
from pourover_amd.solver import HipLBMSolver

pytestmark = pytest.mark.gpu

N = 224


def test_thermal_gaussian_spreading_224():
    print("case: thermal Gaussian 224^3, 100 steps", flush=True)
    tau, steps, sig = 0.8, 100, 6.0
    th = TH.ThermalLBM(grid=(N, N, N))
    th.omega_thermal = 1.0 / tau
    th.tau_thermal = tau
    th.enable_convection_coupling(False)  # pushes omega
    c = torch.arange(N, device="cuda", dtype=torch.float64) - (N - 1) / 2
    r2 = c.view(-1, 1, 1) ** 2 + c.view(1, -1, 1) ** 2 + c.view(1, 1, -1) ** 2
    th._T[th._tcur].copy_(torch.exp(-r2 / (2 * sig ** 2)).float())
    th._init_populations()
    th.is_initialized = True

    def moments():
        T = th._T[th._tcur].double()
        tot = float(T.sum())
        return tot, float((T * c.view(-1, 1, 1) ** 2).sum()) / tot

    tot0, var0 = moments()
    assert th.run(steps)
    tot, var = moments()
    D = 0.25 * (tau - 0.5)
    assert abs(tot - tot0) < 1e-5 * tot0
    assert abs((var - var0) / (2 * D * steps) - 1) < 0.02


def test_planar_interface_stays_planar_224():
    print("case: planar interface 224^3, 20 phase steps", flush=True)
    cfg = SolverConfig(N, N, N, precision="f32", periodic=(False, False, False), force_field=True)
    fluid = HipLBMSolver(cfg)
    fluid.initialize_fields()
    mp = MP.MultiphaseFlow3D(fluid)
    z = torch.arange(N, device="cuda", dtype=torch.float32).view(-1, 1, 1)
    phi = torch.tanh((z - 100.3) / 2.0).expand(N, N, N).contiguous()
    mp._phi[0].copy_(phi)
    mp._phi_new.copy_(phi)
    fluid._macro_valid = True  # u = 0, rho = 1 as initialised
    fluid.ctx.force.zero_()
    mp.accumulate_surface_tension_pre_collision()
    bf = fluid.ctx.force
    assert float(bf[:, 2:-2, 1:-1, 1:-1].abs().max()) == 0.0   # normal (0, 0, 1), div n = 0
    for k in range(20):
        mp.step(12 + k, precollision_applied=True)
    p = mp._phi[mp._cur]
    # the x / y faces keep their phi_new values, so their influence reaches one
    # cell further per step; the columns beyond that cone stay identical
    inner = p[:, 21:-21, 21:-21]
    assert torch.equal(inner, inner[:, :1, :1].expand_as(inner))
    assert float(p[:90].max()) < -0.9 and float(p[110:].min()) > 0.9
