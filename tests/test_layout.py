"""Compact row-run population layout (layout.py): the slot arithmetic the
kernels do from chunk records, emulated in numpy over every non-solid cell
and link, against the slot of the dense source cell."""
from __future__ import annotations

import numpy as np
import pytest

from pourover_amd import _lib as L
from pourover_amd import geometry as G
from pourover_amd.lattice import E, OPP, Q
from pourover_amd.layout import RECORD_WORDS, build_compact_layout


def _lanes(lay, chunk_cells=32):
    """(record row, x, y, z) of every lane of every chunk (lanes past the row end dropped)."""
    rec = lay.records
    base = rec[:, 0].astype(np.int64)
    lane = np.arange(chunk_cells, dtype=np.int64)
    cell = base[:, None] + lane[None, :]
    yz = cell // lay.nx
    x = cell % lay.nx
    ok = (x >= lane[None, :]) & (cell < lay.nx * lay.ny * lay.nz)  # the kernel's wrap test
    r = np.broadcast_to(np.arange(rec.shape[0])[:, None], cell.shape)
    return r[ok], x[ok], (yz % lay.ny)[ok], (yz // lay.ny)[ok]


def _rb(lay, r, a, b):
    return lay.records[r, 1 + a + 3 * b].astype(np.uint32).view(np.int32).astype(np.int64)


def _geometries():
    out = {"v60_48": G.build_v60(G.V60Params.scaled(48))}
    # periodic x/y with a solid slab and a cylinder: wraps plus bounce-back
    n = (40, 36, 32)
    kind = np.zeros((n[2], n[1], n[0]), np.uint8)
    kind[:3] = L.POL_KIND_SOLID
    zz, yy, xx = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    kind[(xx - 20) ** 2 + (yy - 18) ** 2 < 36] = L.POL_KIND_SOLID
    kind[-1] = L.POL_KIND_OUTLET
    flags, rev = G.link_flags(kind, (True, True, False))
    out["periodic_xy"] = (kind, flags, rev, (True, True, False))
    return out


def _unpack(g):
    if isinstance(g, tuple):
        return g
    return g.kind, g.flags, g.rev_flags, (False, False, False)


@pytest.mark.parametrize("name", ["v60_48", "periodic_xy"])
def test_gather_slots_match_dense_sources(name):
    kind, flags, rev, per = _unpack(_geometries()[name])
    nz, ny, nx = kind.shape
    lay = build_compact_layout(flags, per)
    assert lay.records.shape[1] == RECORD_WORDS
    r, x, y, z = _lanes(lay)
    k = kind[z, y, x]
    run = lay.records[r, 10]
    inrun = (x >= (run & 0xFFFF)) & (x < (run >> 16))
    # every non-solid cell is a lane inside its run, and stored exactly once
    act = k != L.POL_KIND_SOLID
    assert inrun[act].all()
    assert act.sum() == (kind != L.POL_KIND_SOLID).sum()
    own = _rb(lay, r, 1, 1) + x
    assert np.array_equal(own[inrun], lay.slot_of(x[inrun], y[inrun], z[inrun]))
    assert own[inrun].min() >= 0 and own[inrun].max() < lay.used_cells
    fl = flags[z, y, x]
    sel = act & (k != L.POL_KIND_INLET)
    for q in range(1, Q):
        ex, ey, ez = (int(v) for v in E[q])
        bounce = ((fl >> np.uint32(q)) & 1).astype(bool)
        # the kernel: source row from the record, x shifted (wrapped or clamped)
        xs = x - ex
        if per[0]:
            xs = xs % nx
        else:
            xs = np.clip(xs, 0, nx - 1)
        slot = _rb(lay, r, 1 - ey, 1 - ez) + xs
        # the dense source the dense gather reads
        ys = y - ey
        ys = ys % ny if per[1] else np.where((ys < 0) | (ys >= ny), y, ys)
        zs = z - ez
        zs = zs % nz if per[2] else np.where((zs < 0) | (zs >= nz), z, zs)
        m = sel & ~bounce
        assert (kind[zs[m], ys[m], xs[m]] != L.POL_KIND_SOLID).all(), q
        assert np.array_equal(slot[m], lay.slot_of(xs[m], ys[m], zs[m])), q


@pytest.mark.parametrize("name", ["v60_48", "periodic_xy"])
def test_reverse_gather_slots(name):
    kind, flags, rev, per = _unpack(_geometries()[name])
    nz, ny, nx = kind.shape
    lay = build_compact_layout(flags, per)
    r, x, y, z = _lanes(lay)
    act = kind[z, y, x] != L.POL_KIND_SOLID
    rf = rev[z, y, x]
    for p in range(Q):
        ex, ey, ez = (int(v) for v in E[p])
        has = act & (((rf >> np.uint32(p)) & 1) == 1)
        xr = (x + ex) % nx if per[0] else x + ex
        slot = _rb(lay, r, 1 + ey, 1 + ez) + xr
        yr = (y + ey) % ny
        zr = (z + ez) % nz
        assert (kind[zr[has], yr[has], xr[has]] != L.POL_KIND_SOLID).all()
        assert np.array_equal(slot[has], lay.slot_of(xr[has], yr[has], zr[has])), p


def test_dense_round_trip_and_every_slot_written():
    g = G.build_v60(G.V60Params.scaled(48))
    lay = build_compact_layout(g.flags)
    rng = np.random.default_rng(3)
    dense = rng.standard_normal((19,) + g.kind.shape)
    comp = lay.from_dense(dense)
    assert comp.shape == (19, lay.stored_cells) and lay.stored_cells % 64 == 0
    back = lay.to_dense(comp, fill=np.nan)
    inrun = np.zeros(g.kind.size, bool)
    inrun[lay.dense_index] = True
    assert np.array_equal(back.reshape(19, -1)[:, inrun], dense.reshape(19, -1)[:, inrun])
    assert np.isnan(back.reshape(19, -1)[:, ~inrun]).all()
    # every used slot is a lane inside its run (written each step)
    r, x, y, z = _lanes(lay)
    run = lay.records[r, 10]
    inr = (x >= (run & 0xFFFF)) & (x < (run >> 16))
    slots = np.sort(lay.slot_of(x[inr], y[inr], z[inr]))
    assert np.array_equal(slots, np.arange(lay.used_cells))
    # fewer slots than the dense chunk lanes
    assert lay.used_cells < 0.9 * lay.n_chunks * 32
