"""Compact row-run population layout.

The dense layout stores populations as [19][nz][ny][nx].  In the V60 box two
thirds of the cells are solid; a launch over 32-cell row chunks still moves
whole 128-byte lines, so the solid tail of every chunk costs HBM traffic.  The
compact layout stores, for each (y, z) row, only the run [x_begin, x_end)
between its first and last non-solid cell, rows one after another in (z, y)
order:

    population q of cell (x, y, z) = plane q, slot row_base[z*ny + y] + x

with plane stride M = stored_cells (the run lengths summed, rounded up to 64
slots so every plane starts on a 256-byte boundary).  Every stored slot is
written each step (non-solid cells with their populations, solid cells inside
a run with zero deviations), so written lines are whole without any solid
slot outside the runs.

A launch uses chunk records instead of the dense chunk list: one record of
RECORD_WORDS uint32 per 32-cell row chunk that meets a run,
    [0]     dense index of the chunk's first cell (x0 + nx (y + ny z))
    [1..9]  row_base of the rows (y + a - 1, z + b - 1), a, b in {0, 1, 2},
            at word 1 + a + 3 b; rows outside a non-periodic face resolve to
            the chunk's own row (the clamp the gather applies), periodic
            faces wrap
    [10]    own row's run: x_begin | x_end << 16
    [11]    0
so a thread finds every neighbour slot from its record without further
lookups: source of link q = word[1 + (1 - e_y) + 3 (1 - e_z)] + x - e_x
(x wrapped or clamped like the dense gather).
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
    used_cells: int          # slots holding a run cell (the rest is padding)
    row_base: np.ndarray     # int64 [nz*ny]
    row_run: np.ndarray      # uint32 [nz*ny]: x_begin | x_end << 16 (0 = empty row)
    records: np.ndarray      # uint32 [n_chunks, RECORD_WORDS]
    dense_index: np.ndarray  # int64 [used_cells]: dense cell of slot s

    @property
    def n_chunks(self) -> int:
        return int(self.records.shape[0])

    def slot_of(self, x, y, z):
        """Slot of dense cell (x, y, z) (valid for cells inside their row's run)."""
        return self.row_base[np.asarray(z) * self.ny + np.asarray(y)] + np.asarray(x)

    def to_dense(self, pop: np.ndarray, fill: float = 0.0) -> np.ndarray:
        """[C, M] stored planes -> [C, nz, ny, nx] (cells outside every run = fill)."""
        c = pop.shape[0]
        out = np.full((c, self.nz * self.ny * self.nx), fill, dtype=pop.dtype)
        out[:, self.dense_index] = pop[:, :self.used_cells]
        return out.reshape(c, self.nz, self.ny, self.nx)

    def from_dense(self, dense: np.ndarray) -> np.ndarray:
        """[C, nz, ny, nx] -> [C, M] (padding slots zero)."""
        c = dense.shape[0]
        flat = dense.reshape(c, -1)
        out = np.zeros((c, self.stored_cells), dtype=dense.dtype)
        out[:, :self.used_cells] = flat[:, self.dense_index]
        return out


def build_compact_layout(flags: np.ndarray, periodic=(False, False, False), chunk_cells: int = 32) -> CompactLayout:
    """Row runs, slot map and chunk records of a cell-flag array [nz, ny, nx]."""
    flags = np.asarray(flags, dtype=np.uint32)
    nz, ny, nx = flags.shape
    if nx >= 1 << 16:
        raise ValueError("the compact layout packs x in 16 bits (nx < 65536)")
    kind = (flags >> np.uint32(L.POL_KIND_SHIFT)) & np.uint32(L.POL_KIND_MASK)
    open_cell = (kind != L.POL_KIND_SOLID).reshape(nz * ny, nx)
    has = open_cell.any(axis=1)
    xb = np.where(has, open_cell.argmax(axis=1), 0).astype(np.int64)
    xe = np.where(has, nx - open_cell[:, ::-1].argmax(axis=1), 0).astype(np.int64)
    length = xe - xb
    start = np.concatenate([[0], np.cumsum(length)[:-1]]).astype(np.int64)
    used = int(length.sum())
    if used == 0:
        raise ValueError("no non-solid cell: nothing to store")
    M = -(-used // PLANE_ALIGN) * PLANE_ALIGN
    if M >= 1 << 31:
        raise ValueError("too many stored cells for 32-bit slots")
    row_base = np.where(has, start - xb, 0).astype(np.int64)
    row_run = (xb | (xe << 16)).astype(np.uint32)
    row_run[~has] = 0
    # dense index of every slot
    rows = np.repeat(np.arange(nz * ny, dtype=np.int64), length)
    offs = np.arange(used, dtype=np.int64) - np.repeat(start, length)
    dense_index = rows * nx + np.repeat(xb, length) + offs
    # chunk list: the chunk_cells-cell row chunks that meet a row's run (so
    # every stored slot is written each step, solid run cells with zeros)
    nch = -(-nx // chunk_cells)
    c0 = np.arange(nch, dtype=np.int64) * chunk_cells
    live = (c0[None, :] < xe[:, None]) & (c0[None, :] + chunk_cells > xb[:, None]) & has[:, None]
    live = live.reshape(nz, ny, nch)
    zz, yy, cc = np.nonzero(live)
    rec = np.zeros((zz.size, RECORD_WORDS), dtype=np.int64)
    rec[:, 0] = (zz * ny + yy) * nx + cc * chunk_cells
    py, pz = bool(periodic[1]), bool(periodic[2])
    for b in range(3):
        z2 = zz + (b - 1)
        z2 = np.where((z2 < 0) | (z2 >= nz), (z2 % nz) if pz else zz, z2)
        for a in range(3):
            y2 = yy + (a - 1)
            y2 = np.where((y2 < 0) | (y2 >= ny), (y2 % ny) if py else yy, y2)
            rec[:, 1 + a + 3 * b] = row_base[z2 * ny + y2]
    rec[:, 10] = row_run[zz * ny + yy]
    records = (rec & 0xFFFFFFFF).astype(np.uint32)  # row bases as two's-complement int32
    return CompactLayout(nx, ny, nz, M, used, row_base, row_run, records, dense_index)
