"""The parity chain closed against the C oracle directly.

* Config-2 size: the 224^3 V60 cup in FP32 storage against the FP64 oracle
  (OpenMP on the host) after 100 steps -- the full-size step is pinned to
  the oracle, not only to the engine's own FP64 run.  Bar: 1e-5 relative L2
  on rho and u over the non-solid cells, the FP32 parity bar of the path.
* 16-bit storage: FP16 and BF16 deviation storage on the 48^3 cup against
  the FP64 oracle after 300 steps.  Bars: FP16 1e-3 relative L2 on u (the
  16-bit storage target), BF16 5e-2 on u (8-bit mantissa; reported), both
  1e-5 on rho.
"""
from __future__ import annotations

import time

import numpy as np
import pytest
import torch

from conftest import rel_l2
from pourover_amd import _lib as L
from pourover_amd import workloads as WL
from pourover_amd.engine import HipContext

pytestmark = pytest.mark.gpu


def _hip(wl, steps):
    ctx = HipContext(wl.config, wl.geometry.flags)
    ctx.init_equilibrium()
    ctx.step(steps)
    red = ctx.macroscopic(macro=True, reduce=True)
    torch.cuda.synchronize()
    out = ctx.rho.double().cpu().numpy(), ctx.u.double().cpu().numpy(), red
    del ctx
    torch.cuda.empty_cache()
    return out


def _oracle(oracle, wl, steps):
    flags = wl.geometry.flags
    n = wl.config.nx
    d = oracle.desc_from_dict(wl.config.oracle_dict(True))
    f = oracle.init_equilibrium(n, n, n)
    for _ in range(steps):
        f, _, _, _ = oracle.step(d, f, flags)
    return oracle.moments(d, f, flags)


def test_v60_224_f32_against_the_oracle(hip, oracle):
    n, steps = 224, 100
    wl = WL.v60_config(n, "f32")
    print("case: v60 224^3 f32 vs FP64 oracle, %d steps" % steps, flush=True)
    rho_g, u_g, red = _hip(wl, steps)
    t0 = time.perf_counter()
    rho_o, u_o, red_o = _oracle(oracle, wl, steps)
    print("oracle %.1f s" % (time.perf_counter() - t0), flush=True)
    fluid = wl.geometry.kind != L.POL_KIND_SOLID
    eu = rel_l2(u_g[:, fluid], u_o[:, fluid])
    er = rel_l2(rho_g[fluid], rho_o[fluid])
    erd = rel_l2(rho_g[fluid] - 1.0, rho_o[fluid] - 1.0)
    print("224^3 f32: rel L2 u %.3e rho %.3e (rho - 1: %.3e)" % (eu, er, erd))
    assert np.abs(u_o[:, fluid]).max() > 1e-3  # the jet and gravity have set the fluid moving
    assert eu < 1e-5
    assert er < 1e-5
    slot = L.RED_SLOTS.index
    assert abs(red[slot("mass")] - red_o[slot("mass")]) <= 1e-7 * red_o[slot("mass")]
    assert abs(red[slot("outlet_flux")] - red_o[slot("outlet_flux")]) <= 1e-4 * abs(red_o[slot("outlet_flux")]) + 1e-12


@pytest.mark.parametrize("precision,u_bar", [("f16", 1e-3), ("bf16", 5e-2)])
def test_v60_16bit_storage_against_the_oracle(hip, oracle, precision, u_bar):
    n, steps = 48, 300
    wl = WL.v60_config(n, precision)
    print("case: v60 48^3 %s vs FP64 oracle, %d steps" % (precision, steps), flush=True)
    rho_g, u_g, _ = _hip(wl, steps)
    rho_o, u_o, _ = _oracle(oracle, WL.v60_config(n, "f64"), steps)
    fluid = wl.geometry.kind != L.POL_KIND_SOLID
    eu = rel_l2(u_g[:, fluid], u_o[:, fluid])
    er = rel_l2(rho_g[fluid], rho_o[fluid])
    print("48^3 %s: rel L2 u %.3e rho %.3e" % (precision, eu, er))
    assert np.isfinite(u_g).all()
    assert eu < u_bar
    assert er < 1e-5
