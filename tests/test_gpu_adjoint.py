"""Adjoint gradients of the outlet flux (csrc/adjoint.hip through the C ABI)
checked against central finite differences of the FP64 forward run and
between the FP32 and FP64 adjoints.

Finite differences need the FP64 forward: one filter cell's K moved by 1e-3
changes J by ~1e-8 relative, below FP32 resolution of J.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import rel_l2
from pourover_amd import workloads as WL
from pourover_amd.adjoint import AdjointProblem, outlet_flux

pytestmark = pytest.mark.gpu


def _problem(n, prec, steps, window, ckpt, les=True):
    wl = WL.v60_config(n, prec, perm_field=True, les=les)
    K = WL.permeability_field(wl.geometry)
    return AdjointProblem(wl, K, steps=steps, window=window, checkpoint=ckpt), K


def _fd_cells(prob, K, cells, eps=1e-3):
    out = []
    for (z, y, x) in cells:
        Kp = K.copy()
        Km = K.copy()
        Kp[z, y, x] *= 1 + eps
        Km[z, y, x] *= 1 - eps
        prob.set_permeability(Kp)
        jp = prob.forward()
        prob.set_permeability(Km)
        jm = prob.forward()
        out.append((jp - jm) / (2 * eps * K[z, y, x]))
    prob.set_permeability(K)
    return np.array(out)


def _pick_cells(geo, k, seed):
    rng = np.random.default_rng(seed)
    por = np.argwhere(geo.porous)
    return [tuple(c) for c in por[rng.choice(len(por), size=k, replace=False)]]


@pytest.mark.parametrize("les", [True, False])
def test_adjoint_matches_finite_differences_small(hip, les):
    """48^3 cup, 60 steps, J over the last 20, checkpoints every 15 (FP64)."""
    prob, K = _problem(48, "f64", 60, 20, 15, les=les)
    J, gK, gu = prob.gradient()
    assert abs(prob.timings["J_adjoint"] - J) <= 1e-12 * abs(J)
    gK = gK.cpu().numpy()
    cells = _pick_cells(prob.wl.geometry, 8, 3)
    fd = _fd_cells(prob, K, cells)
    ad = np.array([gK[c] for c in cells])
    print("K grad adjoint", ad, "fd", fd)
    # FD round-off floor: ~20 steps x 1e-16 x J / (2 eps K) ~ 1e-10 absolute
    assert np.allclose(ad, fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())
    # inlet velocity (z component)
    u0 = prob.ctx.config.inlet_u
    h = 1e-3 * abs(u0[2])
    prob.set_inlet_velocity((u0[0], u0[1], u0[2] + h))
    jp = prob.forward()
    prob.set_inlet_velocity((u0[0], u0[1], u0[2] - h))
    jm = prob.forward()
    prob.set_inlet_velocity(u0)
    fdu = (jp - jm) / (2 * h)
    print("u_in grad adjoint", gu[2], "fd", fdu)
    assert abs(gu[2] - fdu) <= 1e-5 * abs(fdu)
    # gradient is zero outside the porous layer
    assert np.abs(gK[~prob.wl.geometry.porous]).max() == 0.0


def test_adjoint_f32_close_to_f64(hip):
    p64, _ = _problem(48, "f64", 60, 20, 15)
    p32, _ = _problem(48, "f32", 60, 20, 15)
    J64, g64, u64 = p64.gradient()
    J32, g32, u32 = p32.gradient()
    por = p64.wl.geometry.porous
    e = rel_l2(g32.cpu().numpy()[por], g64.cpu().numpy()[por])
    print("J rel %.2e  dJ/dK rel L2 %.2e  dJ/du rel %.2e" % (abs(J32 - J64) / abs(J64), e, abs(u32[2] - u64[2]) / abs(u64[2])))
    assert abs(J32 - J64) <= 1e-4 * abs(J64)
    assert e < 1e-3
    assert abs(u32[2] - u64[2]) <= 1e-3 * abs(u64[2])


def test_store_all_equals_checkpointing(hip):
    """checkpoint=None keeps every forward state; the reverse sweep then
    recomputes nothing and the gradient is the checkpointed one, bit for bit.
    The checkpointed forward run leaves its last segment in the segment store
    (the sweep rebuilds only the others); every forward form ends with the
    final state as the context's current one, equal to a plain run's."""
    pc, K = _problem(48, "f64", 60, 20, 15)
    ps, _ = _problem(48, "f64", 60, 20, None)
    Jc, gc, uc = pc.gradient()
    final_c = pc.ctx.f_current.clone()
    Js, gs, us = ps.gradient()
    final_s = ps.ctx.f_current.clone()
    assert Jc == Js
    assert torch.equal(gc, gs)
    assert np.array_equal(uc, us)
    pc.forward(checkpoints=False)
    assert torch.equal(pc.ctx.f_current, final_c) and torch.equal(final_s, final_c)
    # the last segment the forward run stored is the one the sweep would rebuild
    pc.forward(checkpoints=True)
    held = pc.seg.clone()
    pc._rebuild(pc.steps // pc.checkpoint - 1, pc.seg)
    assert torch.equal(pc.seg, held)


def test_compact_layout_adjoint_equals_dense(hip):
    """The row-run population layout changes where populations live, not the
    arithmetic: the per-cell gradient is the dense layout's, bit for bit."""
    wl = WL.v60_config(48, "f64", perm_field=True)
    K = WL.permeability_field(wl.geometry)
    pd = AdjointProblem(wl, K, steps=40, window=10, checkpoint=10, compact=False)
    pc = AdjointProblem(wl, K, steps=40, window=10, checkpoint=10, compact=True)
    assert pc.ctx.plane < pd.ctx.plane
    Jd, gd, ud = pd.gradient()
    Jc, gc, uc = pc.gradient()
    # J and dJ/du_in are FP64 block sums, added in launch order (chunk sizes differ)
    assert abs(Jd - Jc) <= 1e-12 * abs(Jd)
    assert torch.equal(gd, gc)
    assert np.allclose(ud, uc, rtol=1e-12, atol=0)


def test_autograd_function(hip):
    prob, K = _problem(32, "f64", 30, 10, 10)
    Kt = torch.tensor(K, requires_grad=True)
    ut = torch.tensor(prob.ctx.config.inlet_u, dtype=torch.float64, requires_grad=True)
    J = outlet_flux(Kt, ut, prob)
    (2.0 * J).backward()
    _, gK, gu = prob.gradient()
    assert torch.allclose(Kt.grad, 2.0 * gK.cpu(), rtol=1e-12, atol=0)
    assert np.allclose(ut.grad.numpy(), 2.0 * gu, rtol=1e-12)


def test_config5_adjoint_128(hip):
    """128^3 scaled cup, 500 steps, J over the last 100, checkpoints every 50:
    FP64 adjoint vs central FD (eps 1e-3) on 16 random filter cells and on
    u_in (1e-3 relative); FP32 adjoint vs FP64 adjoint (1e-3 relative L2 on
    dJ/dK, 1e-3 on dJ/du_in); reverse cost per forward step."""
    p64, K = _problem(128, "f64", 500, 100, 50)
    J, gK, gu = p64.gradient()
    gK = gK.cpu().numpy()
    cells = _pick_cells(p64.wl.geometry, 16, 7)
    fd = _fd_cells(p64, K, cells)
    ad = np.array([gK[c] for c in cells])
    rel = np.abs(ad - fd) / np.maximum(np.abs(fd), 1e-300)
    print("config5 J %.6e max rel err vs FD %.2e" % (J, rel.max()))
    assert rel.max() < 1e-3
    # inlet velocity: central FD of the FP64 forward run (eps 1e-3 relative)
    u0 = p64.ctx.config.inlet_u
    h = 1e-3 * abs(u0[2])
    p64.set_inlet_velocity((u0[0], u0[1], u0[2] + h))
    jp = p64.forward()
    p64.set_inlet_velocity((u0[0], u0[1], u0[2] - h))
    jm = p64.forward()
    p64.set_inlet_velocity(u0)
    fdu = (jp - jm) / (2 * h)
    print("config5 u_in grad adjoint %.6e fd %.6e" % (gu[2], fdu))
    assert fdu != 0.0
    assert abs(gu[2] - fdu) <= 1e-3 * abs(fdu)
    p32, _ = _problem(128, "f32", 500, 100, 50)
    J32, g32, u32 = p32.gradient()
    t_fwd = p32.forward_time()
    step_cost = p32.timings["reverse_s"] / t_fwd   # recompute + adjoint kernel per forward step
    por = p64.wl.geometry.porous
    e = rel_l2(g32.cpu().numpy()[por], gK[por])
    eu = abs(u32[2] - gu[2]) / abs(gu[2])
    print("config5 f32 vs f64 dJ/dK rel L2 %.2e, dJ/du_in rel %.2e, adjoint step / forward step %.2f, "
          "whole gradient / forward %.2f" % (e, eu, step_cost, (p32.timings["forward_s"] + p32.timings["reverse_s"]) / t_fwd))
    assert e < 1e-3
    assert eu < 1e-3
