"""pol_reduce_noalloc, the allocation-free form of pol_reduce, through the C ABI
(collected last: it shares the static scratch with nothing else, and a
failure here cannot hide the tests before it).

Its FP64 reductions of a state must equal pol_reduce's on the same state bit
for bit (same kernels, same block partition, the partials in the static
scratch instead of the context's buffer), on two calls in a row whose second
state is smaller (fewer blocks: nothing of the first call's partials may
leak into the second)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from pourover_amd import _lib as L
from pourover_amd import workloads as WL
from pourover_amd.engine import HipContext

pytestmark = pytest.mark.gpu


def _state(kind):
    if kind == "v60_48":
        wl = WL.v60_config(48, "f32")
        ctx = HipContext(wl.config, wl.geometry.flags)
        ctx.init_equilibrium()
        ctx.step(37)
    else:
        cfg = WL.taylor_green_config(24, "f32")
        rho0, u0 = WL.taylor_green_fields(24)
        ctx = HipContext(cfg)
        ctx.init_equilibrium(torch.from_numpy(rho0), torch.from_numpy(u0))
        ctx.step(11)
    torch.cuda.synchronize()
    return ctx


def _noalloc(ctx):
    b = ctx.buffers()
    b.f[0] = ctx.f_current.data_ptr()
    b.f[1] = ctx.f_next.data_ptr()
    out = torch.full((L.POL_NRED,), float("nan"), dtype=torch.float64, device=ctx.device)
    desc = ctx.config.desc(ctx.flags is not None)
    st = ctx.lib.pol_reduce_noalloc(C.byref(desc), C.byref(b), L.ptr(out), L.stream_handle())
    assert st == L.POL_OK, ctx.lib.pol_last_error(None)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_reduction_equals_pol_reduce_twice_second_smaller(hip):
    got = []
    for kind in ("v60_48", "tg_24"):
        print("case: pol_reduce_noalloc on %s" % kind, flush=True)
        ctx = _state(kind)
        want = ctx.macroscopic(macro=False, reduce=True)
        have = _noalloc(ctx)
        assert np.isfinite(have).all()
        assert np.array_equal(have.view(np.uint64), np.asarray(want, dtype=np.float64).view(np.uint64)), (kind, have, want)
        got.append(have)
    # the two states differ (the second call did not return the first's result)
    assert got[0][L.RED_SLOTS.index("cells")] != got[1][L.RED_SLOTS.index("cells")]
