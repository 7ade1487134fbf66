"""pol_stream_only: collisionless streaming -- the step's gather, bounce-back
links, open faces and stores without its collision (the layout's bandwidth
ceiling that bench.py times beside the step).

- a periodic 32^3 box of random deviations: one pass equals numpy's roll of
  each direction by its lattice vector, bit for bit (FP32, FP16);
- the 48^3 V60 cup in the compact layout: each of three passes equals a numpy
  statement of the streaming rules applied to the previous pass (pull from
  x - e_q with every coordinate clamped at the box faces, a bounce-back link
  reads the cell's own opposite population, inlet cells gather nothing) on
  every non-solid cell, bit for bit (FP32, FP16, BF16).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from pourover_amd import _lib as L
from pourover_amd import workloads as WL
from pourover_amd.engine import HipContext
from pourover_amd.lattice import E, OPP

pytestmark = pytest.mark.gpu

TORCH = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}


def _bits(t: torch.Tensor) -> np.ndarray:
    a = t.cpu()
    return a.view(torch.int16).numpy() if a.element_size() == 2 else a.view(torch.int32).numpy()


@pytest.mark.parametrize("precision", ["f32", "f16"])
def test_periodic_box_is_a_roll(hip, precision):
    n = 32
    print("case: stream_only periodic %d^3 %s" % (n, precision), flush=True)
    ctx = HipContext(WL.taylor_green_config(n, precision))
    g = torch.Generator().manual_seed(5)
    f0 = (torch.randn((19, n, n, n), generator=g) * 0.5).to(TORCH[precision])
    ctx.load_dense_populations(f0.cuda())
    ctx.stream_only(1)
    got = ctx.dense_populations().cpu()
    for q in range(19):
        ex, ey, ez = (int(v) for v in E[q])
        want = torch.roll(f0[q], shifts=(ez, ey, ex), dims=(0, 1, 2))
        assert np.array_equal(_bits(got[q]), _bits(want)), "direction %d" % q


def _stream_numpy(f: np.ndarray, flags: np.ndarray) -> np.ndarray:
    """One pull-streaming pass of dense populations [19, nz, ny, nx] under the
    cell flags: the rules of the fused step's gather, without the collision."""
    nz, ny, nx = flags.shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    kind = (flags >> L.POL_KIND_SHIFT) & L.POL_KIND_MASK
    out = np.empty_like(f)
    for q in range(19):
        ex, ey, ez = (int(v) for v in E[q])
        src = f[q][np.clip(z - ez, 0, nz - 1), np.clip(y - ey, 0, ny - 1), np.clip(x - ex, 0, nx - 1)]
        if q:
            bounce = (flags >> q) & 1 == 1
            src = np.where(bounce, f[OPP[q]], src)
        out[q] = np.where(kind == L.POL_KIND_INLET, np.zeros_like(src), src)
    return out


@pytest.mark.parametrize("precision", ["f32", "f16", "bf16"])
def test_v60_compact_streaming_rules(hip, precision):
    n = 48
    wl = WL.v60_config(n, precision)
    flags = wl.geometry.flags
    print("case: stream_only V60 %d^3 %s compact" % (n, precision), flush=True)
    ctx = HipContext(wl.config, flags)
    assert ctx.compact
    g = torch.Generator().manual_seed(9)
    dense = (torch.randn((19, n, n, n), generator=g) * 0.5).to(TORCH[precision])
    ctx.load_dense_populations(dense.cuda())
    fluid = wl.geometry.kind != L.POL_KIND_SOLID
    prev = ctx.dense_populations().cpu()
    for p in range(3):
        ctx.stream_only(1)
        got = ctx.dense_populations().cpu()
        want = torch.from_numpy(_stream_numpy(prev.float().numpy(), flags)).to(TORCH[precision])
        gb, wb = _bits(got), _bits(want)
        bad = (gb != wb) & fluid[None]
        assert not bad.any(), "pass %d: %d populations differ" % (p, int(bad.sum()))
        prev = got
