"""Compact row-run population layout.

The dense layout stores populations as [19][nz][ny][nx].  In the V60 box two
thirds of the cells are solid; a launch over 32-cell row chunks still moves
whole 128-byte lines, so the solid tail of every chunk costs HBM traffic.  The
compact layout stores, for each (y, z) row, only the run [x_begin, x_end)
between its first and last non-solid cell, rows one after another in (z, y)
order:

    population q of cell (x, y, z) = plane q, slot row_base[z*ny + y] + x

with plane stride M = stored_cells (a leading guard of 64 slots,
the run lengths summed, at least two trailing guard slots, rounded up to 64
slots so every plane starts on a 256-byte boundary).  Every stored slot of a
run is written each step (non-solid cells with their populations, solid
cells inside a run with zero deviations), so written lines are whole without
any solid slot outside the runs.  The guards keep slots 0, 1, M - 2 and
M - 1 free of cells: a kernel may then clamp a row offset (source slot + e_x)
into [1, M - 2] once and reach the x - 1 / x + 1 neighbours with constant
offsets, without the clamp ever moving a cell's own source.

A launch lists its row pieces (`pieces`: the dense index x0 + nx (y + ny z)
of each piece's first cell, one uint32 per chunk of chunk_cells cells that
meets a run) and reads one row record of RECORD_WORDS uint32 per dense row
(`row_records[z * ny + y]`, the row found from the piece's first cell by one
division by nx):
    [0]     dense index of the row's first cell (nx (y + ny z))
    [1..9]  row_base of the rows (y + a - 1, z + b - 1), a, b in {0, 1, 2},
            at word 1 + a + 3 b; rows outside a non-periodic face resolve to
            the row itself (the clamp the gather applies), periodic faces
            wrap
    [10]    the row's run: x_begin | x_end << 16
    [11]    the row: y | z << 16
so a thread finds every neighbour slot from its row record without further
lookups: source of link q = word[1 + (1 - e_y) + 3 (1 - e_z)] + x - e_x
(x wrapped or clamped like the dense gather).  The cell index and its x come
from the piece alone, so the cell's flags and the row record load side by
side.  The device piece words (`piece_words`) carry bit 31 for a piece all
of whose cells inside the row hold the flag word 0 (fluid, no bounce link,
not porous, LES on): the step kernel takes that word without loading it,
which skips most of the flag array's 4 B per cell (85 % of the 4-cell
pieces of the 224^3 cup).  Pieces of one row share its record: 4 B per piece plus 48 B per row
(about 4 MB at 224^3) instead of a 48-byte record per piece (23 MB).  With
align > 1 each row's storage starts on an aligned slot and is padded to a
multiple of align, and the pieces start at the run's first cell (see
build_compact_layout).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib as L

RECORD_WORDS = 12
PLANE_ALIGN = 64  # slots


@dataclass
class CompactLayout:
    nx: int
    ny: int
    nz: int
    stored_cells: int        # M: plane stride in slots
    used_cells: int          # leading guard + slots of the stored runs (padding included)
    row_base: np.ndarray     # int64 [nz*ny]
    row_run: np.ndarray      # uint32 [nz*ny]: x_begin | x_end << 16 (0 = empty row); x_end may pass nx (padding)
    pieces: np.ndarray       # uint32 [n_chunks]: dense index of each launched piece's first cell
    row_records: np.ndarray  # uint32 [nz*ny, RECORD_WORDS]
    plain: np.ndarray        # bool [n_chunks]: every cell of the piece inside the row has flags 0
    slot_dense: np.ndarray   # int64 [used_cells]: dense cell of slot s, -1 for guard and padding
    align: int = 1           # slots a row's storage is aligned and padded to
    x_face: bool = True      # some run holds a cell at x = 0 or x = nx - 1

    @property
    def n_chunks(self) -> int:
        return int(self.pieces.shape[0])

    @property
    def piece_words(self) -> np.ndarray:
        """The launched piece list as the kernels take it: the first cell, with
        bit 31 set for a plain piece (its cells' flag words are 0)."""
        return self.pieces | np.where(self.plain, np.uint32(L.POL_PIECE_PLAIN), np.uint32(0))

    @property
    def records(self) -> np.ndarray:
        """Per-piece view [n_chunks, RECORD_WORDS] of what a kernel decodes for
        a piece: word 0 the piece's first cell, words 1..11 its row record."""
        rec = self.row_records[self.pieces.astype(np.int64) // self.nx].copy()
        rec[:, 0] = self.pieces
        return rec

    @property
    def dense_index(self) -> np.ndarray:
        """Dense cells of the run slots in slot order (padding left out)."""
        return self.slot_dense[self.slot_dense >= 0]

    @property
    def run_slots(self) -> np.ndarray:
        """Slots that hold a run cell (the slots dense_index maps)."""
        return np.nonzero(self.slot_dense >= 0)[0]

    def slot_of(self, x, y, z):
        """Slot of dense cell (x, y, z) (valid for cells inside their row's run)."""
        return self.row_base[np.asarray(z) * self.ny + np.asarray(y)] + np.asarray(x)

    def to_dense(self, pop: np.ndarray, fill: float = 0.0) -> np.ndarray:
        """[C, M] stored planes -> [C, nz, ny, nx] (cells outside every run = fill)."""
        c = pop.shape[0]
        out = np.full((c, self.nz * self.ny * self.nx), fill, dtype=pop.dtype)
        out[:, self.dense_index] = pop[:, self.run_slots]
        return out.reshape(c, self.nz, self.ny, self.nx)

    def from_dense(self, dense: np.ndarray) -> np.ndarray:
        """[C, nz, ny, nx] -> [C, M] (padding slots zero)."""
        c = dense.shape[0]
        flat = dense.reshape(c, -1)
        out = np.zeros((c, self.stored_cells), dtype=dense.dtype)
        out[:, self.run_slots] = flat[:, self.dense_index]
        return out


def build_compact_layout(flags: np.ndarray, periodic=(False, False, False), chunk_cells: int = 32,
                         align: int = 1) -> CompactLayout:
    """Row runs, slot map, launched pieces and row records of a cell-flag
    array [nz, ny, nx].

    align = 1 packs the runs back to back and launches the chunk_cells-cell
    row chunks (x0 a multiple of chunk_cells) that meet a run.  align > 1
    (dividing chunk_cells) starts every row's storage on a multiple of align
    slots, pads it to a multiple of align, and launches run pieces of
    chunk_cells cells from the run's first cell, so every launched chunk
    writes whole aligned slot segments (padding slots take zero deviations).
    """
    flags = np.asarray(flags, dtype=np.uint32)
    nz, ny, nx = flags.shape
    if chunk_cells % align:
        raise ValueError("align must divide chunk_cells")
    if nx + align >= 1 << 16 or ny >= 1 << 16 or nz >= 1 << 16:
        raise ValueError("the compact layout packs x, y and z in 16 bits (below 65536)")
    kind = (flags >> np.uint32(L.POL_KIND_SHIFT)) & np.uint32(L.POL_KIND_MASK)
    open_cell = (kind != L.POL_KIND_SOLID).reshape(nz * ny, nx)
    has = open_cell.any(axis=1)
    xb = np.where(has, open_cell.argmax(axis=1), 0).astype(np.int64)
    xe = np.where(has, nx - open_cell[:, ::-1].argmax(axis=1), 0).astype(np.int64)
    length = xe - xb
    padded = -(-length // align) * align
    # a whole PLANE_ALIGN block: the guard shifts no row off its 128-byte
    # line parity (a shift by half a line costs each wave a third line)
    guard = -(-max(align, 2) // PLANE_ALIGN) * PLANE_ALIGN
    start = guard + np.concatenate([[0], np.cumsum(padded)[:-1]]).astype(np.int64)
    if int(padded.sum()) == 0:
        raise ValueError("no non-solid cell: nothing to store")
    used = guard + int(padded.sum())
    M = -(-(used + 2) // PLANE_ALIGN) * PLANE_ALIGN
    if M >= 1 << 31:
        raise ValueError("too many stored cells for 32-bit slots")
    row_base = np.where(has, start - xb, 0).astype(np.int64)
    row_run = (xb | ((xb + padded) << 16)).astype(np.uint32)
    row_run[~has] = 0
    # dense cell of every slot (-1 for guard and padding slots)
    rows = np.repeat(np.arange(nz * ny, dtype=np.int64), padded)
    offs = np.arange(guard, used, dtype=np.int64) - np.repeat(start, padded)
    slot_dense = np.concatenate([np.full(guard, -1, np.int64),
                                 np.where(offs < np.repeat(length, padded), rows * nx + np.repeat(xb, padded) + offs, -1)])
    if align == 1:
        # chunk_cells-cell row chunks (x0 a multiple of chunk_cells) that meet a
        # row's run (so every stored slot is written each step)
        nch = -(-nx // chunk_cells)
        c0 = np.arange(nch, dtype=np.int64) * chunk_cells
        live = (c0[None, :] < xe[:, None]) & (c0[None, :] + chunk_cells > xb[:, None]) & has[:, None]
        rr, cc = np.nonzero(live)
        first = rr * nx + cc * chunk_cells
    else:
        # run pieces of chunk_cells cells from each run's first cell
        pieces = -(-padded // chunk_cells)
        rr = np.repeat(np.arange(nz * ny, dtype=np.int64), pieces)
        k = np.arange(int(pieces.sum()), dtype=np.int64) - np.repeat(np.cumsum(pieces) - pieces, pieces)
        first = rr * nx + xb[rr] + k * chunk_cells
    if nx * ny * nz >= 1 << 31:
        raise ValueError("dense cell indices must fit 31 bits (bit 31 of a piece word marks a plain piece)")
    pieces = first.astype(np.uint32)
    # plain pieces: every lane inside the row (x < nx) holds flag word 0
    lane = np.arange(chunk_cells, dtype=np.int64)
    x0 = first % nx
    fl = flags.reshape(-1)
    cells = np.minimum(first[:, None] + lane[None, :], nx * ny * nz - 1)
    inside = x0[:, None] + lane[None, :] < nx
    plain = ~(inside & (fl[cells] != 0)).any(axis=1)
    # one record per dense row (rows without a run keep theirs unreferenced)
    ra = np.arange(nz * ny, dtype=np.int64)
    zz, yy = ra // ny, ra % ny
    rec = np.zeros((nz * ny, RECORD_WORDS), dtype=np.int64)
    rec[:, 0] = ra * nx
    py, pz = bool(periodic[1]), bool(periodic[2])
    for b in range(3):
        z2 = zz + (b - 1)
        z2 = np.where((z2 < 0) | (z2 >= nz), (z2 % nz) if pz else zz, z2)
        for a in range(3):
            y2 = yy + (a - 1)
            y2 = np.where((y2 < 0) | (y2 >= ny), (y2 % ny) if py else yy, y2)
            rec[:, 1 + a + 3 * b] = np.where(has, row_base[z2 * ny + y2], 0)
    rec[:, 10] = row_run
    rec[:, 11] = yy | (zz << 16)
    row_records = (rec & 0xFFFFFFFF).astype(np.uint32)  # row bases as two's-complement int32
    x_face = bool((has & ((xb == 0) | (xe == nx))).any())
    return CompactLayout(nx, ny, nz, M, used, row_base, row_run, pieces, row_records, plain, slot_dense, align, x_face)
