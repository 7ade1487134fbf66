"""Compact row-run population layout (layout.py): the slot arithmetic the
kernels do from chunk records, emulated in numpy over every non-solid cell
and link, against the slot of the dense source cell."""
from __future__ import annotations

import numpy as np
import pytest

from pourover_amd import _lib as L
from pourover_amd import geometry as G
from pourover_amd.lattice import E, Q
from pourover_amd.layout import RECORD_WORDS, build_compact_layout


def _lanes(lay, chunk_cells=32, with_padding=False):
    """(piece, x, y, z) of every lane of every piece; lanes past the row end
    dropped (the kernel's wrap test), or with with_padding also returned
    with their unwrapped x (piece, sx, y, z, valid)."""
    # the kernels' decode (lbm_common.h item_record): the piece word's first
    # cell (bit 31 masked off), its row = first // nx selects the row record,
    # x0 = first - row nx, lanes at x0 + lane, valid below nx; y, z from the
    # record's word 11
    words = lay.piece_words
    assert words.dtype == np.uint32
    base = (words & np.uint32(~L.POL_PIECE_PLAIN & 0xFFFFFFFF)).astype(np.int64)
    assert np.array_equal(base, lay.pieces.astype(np.int64))
    row = base // lay.nx
    rec = lay.row_records[row]
    assert np.array_equal(lay.records[:, 1:], rec[:, 1:]) and np.array_equal(lay.records[:, 0], lay.pieces)
    y0 = (rec[:, 11] & 0xFFFF).astype(np.int64)
    z0 = (rec[:, 11] >> 16).astype(np.int64)
    assert np.array_equal(z0 * lay.ny + y0, row)  # word 11 is the row's (y, z)
    assert np.array_equal(rec[:, 0].astype(np.int64), row * lay.nx)  # word 0 its first cell
    x0 = base - row * lay.nx
    lane = np.arange(chunk_cells, dtype=np.int64)
    cell = base[:, None] + lane[None, :]
    sx = x0[:, None] + lane[None, :]
    ok = (sx < lay.nx) & (cell < lay.nx * lay.ny * lay.nz)
    r = np.broadcast_to(np.arange(rec.shape[0])[:, None], cell.shape)
    yb = np.broadcast_to(y0[:, None], cell.shape)
    zb = np.broadcast_to(z0[:, None], cell.shape)
    if with_padding:
        return r.ravel(), sx.ravel(), yb.ravel(), zb.ravel(), ok.ravel()
    return r[ok], sx[ok], yb[ok], zb[ok]


LAYOUTS = [(1, 32), (4, 4), (8, 8), (32, 32), (16, 16), (32, 64), (64, 64)]


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


@pytest.mark.parametrize("align,chunk", LAYOUTS)
@pytest.mark.parametrize("name", ["v60_48", "periodic_xy"])
def test_gather_slots_match_dense_sources(name, align, chunk):
    kind, flags, rev, per = _unpack(_geometries()[name])
    nz, ny, nx = kind.shape
    lay = build_compact_layout(flags, per, chunk_cells=chunk, align=align)
    assert lay.row_records.shape == (nz * ny, RECORD_WORDS) and lay.pieces.shape == (lay.n_chunks,)
    r, x, y, z = _lanes(lay, chunk)
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
        # the kernel's fast form for lanes off the x faces (CompactRows): the
        # row offset bounded once by an unsigned minimum, ro = min(u32(rb + x
        # - 1), M - 3), the load at ro + 1 - e_x; every link (bounce links
        # too, whose value is replaced) stays inside the plane
        fast = sel & (x > 0) & (x < nx - 1)
        M = lay.stored_cells
        ro = np.minimum((_rb(lay, r, 1 - ey, 1 - ez) + x - 1) & 0xFFFFFFFF, M - 3)
        idx = ro + 1 - ex
        assert np.array_equal(idx[fast & ~bounce], slot[fast & ~bounce]), q
        assert idx.min() >= 0 and idx.max() <= M - 1


@pytest.mark.parametrize("align,chunk", LAYOUTS)
@pytest.mark.parametrize("name", ["v60_48", "periodic_xy"])
def test_reverse_gather_slots(name, align, chunk):
    kind, flags, rev, per = _unpack(_geometries()[name])
    nz, ny, nx = kind.shape
    lay = build_compact_layout(flags, per, chunk_cells=chunk, align=align)
    r, x, y, z = _lanes(lay, chunk)
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


@pytest.mark.parametrize("align,chunk", LAYOUTS)
def test_dense_round_trip_and_every_slot_written(align, chunk):
    g = G.build_v60(G.V60Params.scaled(48))
    lay = build_compact_layout(g.flags, chunk_cells=chunk, align=align)
    rng = np.random.default_rng(3)
    dense = rng.standard_normal((19,) + g.kind.shape)
    comp = lay.from_dense(dense)
    assert comp.shape == (19, lay.stored_cells) and lay.stored_cells % 64 == 0
    back = lay.to_dense(comp, fill=np.nan)
    inrun = np.zeros(g.kind.size, bool)
    inrun[lay.dense_index] = True
    assert inrun.sum() == (g.kind != L.POL_KIND_SOLID).sum()
    assert np.array_equal(back.reshape(19, -1)[:, inrun], dense.reshape(19, -1)[:, inrun])
    assert np.isnan(back.reshape(19, -1)[:, ~inrun]).all()
    # every used slot (padding included) is owned by exactly one lane
    r, sx, y, z, ok = _lanes(lay, chunk, with_padding=True)
    run = lay.records[r, 10].astype(np.int64)
    inr = (sx >= (run & 0xFFFF)) & (sx < (run >> 16))
    slots = np.sort(lay.row_base[z[inr] * lay.ny + y[inr]] + sx[inr])
    guard = max(64, align)
    assert np.array_equal(slots, np.arange(guard, lay.used_cells))
    # guard slots at both plane ends hold no cell
    assert (lay.slot_dense[:guard] < 0).all() and lay.used_cells <= lay.stored_cells - 2
    # a padding lane is never a valid cell of the grid
    pad = inr & (lay.slot_dense[np.clip(lay.row_base[z * lay.ny + y] + sx, 0, lay.used_cells - 1)] < 0)
    assert not (pad & ok & (sx < lay.nx) & (g.kind.reshape(-1)[np.clip((z * lay.ny + y) * lay.nx + sx, 0, g.kind.size - 1)]
                                            != L.POL_KIND_SOLID)).any()
    if align > 1:
        # every row's storage starts aligned and launched pieces write whole segments
        starts = (lay.row_base + (lay.row_run & 0xFFFF))[lay.row_run != 0]
        assert (starts % align == 0).all()
        assert lay.used_cells % align == 0
    if align == 1:
        # packed runs: one slot per run cell
        assert lay.used_cells == 64 + int(((lay.row_run >> 16) - (lay.row_run & 0xFFFF)).astype(np.int64).sum())


@pytest.mark.parametrize("align,chunk", LAYOUTS)
@pytest.mark.parametrize("name", ["v60_48", "periodic_xy"])
def test_plain_pieces(name, align, chunk):
    """A piece word carries POL_PIECE_PLAIN exactly when every lane of the
    piece inside its row has flag word 0: the kernels then take 0 for those
    lanes without loading the flags, so a set bit with a nonzero flag under
    it would run a wall or porous cell as plain fluid."""
    kind, flags, rev, per = _unpack(_geometries()[name])
    nz, ny, nx = kind.shape
    lay = build_compact_layout(flags, per, chunk_cells=chunk, align=align)
    words = lay.piece_words
    plain = (words & np.uint32(L.POL_PIECE_PLAIN)) != 0
    assert np.array_equal(plain, lay.plain)
    lane = np.arange(chunk, dtype=np.int64)
    first = lay.pieces.astype(np.int64)
    inside = (first % nx)[:, None] + lane[None, :] < nx
    fl = flags.reshape(-1)[np.minimum(first[:, None] + lane[None, :], flags.size - 1)]
    zero = np.where(inside, fl == 0, True).all(axis=1)
    assert np.array_equal(plain, zero)
    # wall pieces are never plain; short pieces of the interior are
    assert not plain.all()
    if chunk <= 8:
        assert plain.any()


def test_x_face_flag():
    """x_face is set exactly when some run holds a cell on an x face (the
    kernels need the x wrap / clamp variant of the gather then)."""
    g = G.build_v60(G.V60Params.scaled(48))
    lay = build_compact_layout(g.flags, chunk_cells=16, align=16)
    open_cell = g.kind != L.POL_KIND_SOLID
    assert lay.x_face == bool(open_cell[:, :, 0].any() or open_cell[:, :, -1].any())
    kind, flags, rev, per = _geometries()["periodic_xy"]
    assert build_compact_layout(flags, per, chunk_cells=16, align=16).x_face
