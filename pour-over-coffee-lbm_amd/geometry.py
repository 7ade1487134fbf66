"""V60 brewer geometry and the per-cell flag words the kernels read.

The reference builds its masks in Taichi kernels (filter_paper.py:206-286 for
the cone / domain shell, :288-364 for the filter zone) from the physical cup
dimensions.  The workloads here fix the geometry in lattice units instead
(30 deg cone half-angle, outlet hole radius 8 cells, 2-cell filter layer,
inlet jet radius 4 cells at 224^3), built on the host once, with integer
results (bit-exact, pinned by hashes in tests/golden).

Layout of one uint32 flag word per cell (include/pourover_hip.h):
  bits 0..18  link q bounces (its source x - e_q is a wall)
  bits 24..26 kind: 0 fluid, 1 solid, 2 inlet, 3 outlet
  bit 27      porous (Darcy-Forchheimer, LES off)
  bit 28      LES off
Arrays are indexed [z, y, x] (x fastest), matching the device layout.
"""
from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass, asdict

import numpy as np

from . import _lib as L
from .lattice import E, Q


@dataclass(frozen=True)
class V60Params:
    """Lattice-unit V60 description (defaults: the 224^3 brew workload)."""

    n: int = 224
    half_angle_deg: float = 30.0
    hole_radius: float = 8.0
    hole_length: int = 8          # cells below the cone base (z = 0 .. hole_length-1)
    filter_thickness: float = 2.0
    gap: float = 3.0              # paper-to-wall drainage gap (~2 mm, filter_paper.py:247)
    shell: int = 3                # solid domain shell on x/y faces (filter_paper.py:277-280)
    wall: float = 2.0             # wall thickness inside the shell
    inlet_radius: float = 4.0

    @staticmethod
    def scaled(n: int) -> "V60Params":
        """The 224^3 geometry scaled by n/224 (radii, jet, layer, gap), with floors that
        keep every feature resolved on small grids (hole radius >= 3, jet radius >= 2,
        filter >= 2 cells -- thinner discrete shells let diagonal D3Q19 links bypass
        the paper -- gap >= 1.5, wall >= 1, hole length >= 3)."""
        s = n / 224.0
        return V60Params(n=n, hole_radius=max(3.0, 8.0 * s), hole_length=max(3, int(round(8 * s))),
                         filter_thickness=max(2.0, 2.0 * s), gap=max(1.5, 3.0 * s), shell=3,
                         wall=max(1.0, 2.0 * s), inlet_radius=max(2.0, 4.0 * s))


@dataclass
class Geometry:
    params: V60Params
    kind: np.ndarray      # uint8 [n,n,n]
    porous: np.ndarray    # bool
    flags: np.ndarray     # uint32
    rev_flags: np.ndarray  # uint32 (adjoint reverse gather)
    cone_top_z: int

    def counts(self) -> dict:
        k = self.kind
        return {
            "fluid": int((k == L.POL_KIND_FLUID).sum()),
            "solid": int((k == L.POL_KIND_SOLID).sum()),
            "inlet": int((k == L.POL_KIND_INLET).sum()),
            "outlet": int((k == L.POL_KIND_OUTLET).sum()),
            "porous": int(self.porous.sum()),
        }

    def digest(self) -> str:
        h = hashlib.sha256()
        h.update(self.flags.tobytes())
        return h.hexdigest()


def build_v60(p: V60Params = V60Params()) -> Geometry:
    n = p.n
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    xc = yc = (n - 1) / 2.0
    r = np.sqrt((x - xc) ** 2 + (y - yc) ** 2)
    tan = math.tan(math.radians(p.half_angle_deg))
    r_max = (n - 1) / 2.0 - p.shell - p.wall          # largest open radius
    zb = p.hole_length                                 # cone base plane
    r_wall = np.minimum(p.hole_radius + (z - zb) * tan, r_max)   # open radius of the cup
    r_paper = r_wall - p.gap                           # outer surface of the paper
    cone_top = zb + int(math.ceil((r_max - p.hole_radius) / tan))

    solid = np.zeros((n, n, n), dtype=bool)
    below = z < zb
    solid |= below & (r > p.hole_radius)               # base plate around the hole
    solid |= (~below) & (r > r_wall)                   # cup wall and everything outside
    # domain shell on the x/y faces, lid on the top plane, floor on z = 0
    solid |= (x < p.shell) | (x > n - 1 - p.shell) | (y < p.shell) | (y > n - 1 - p.shell)
    top = z == n - 1
    inlet = top & (r <= p.inlet_radius)
    solid |= top & ~inlet
    outlet = (z == 0) & ~solid

    # filter paper: shell of thickness t inside r_paper, closed by a disk at the base
    t = p.filter_thickness
    above = (z >= zb) & (z < n - 1)
    paper = above & (r <= r_paper) & (r > r_paper - t)
    paper |= (z >= zb) & (z < zb + t) & (r <= r_paper)
    paper &= ~solid

    kind = np.full((n, n, n), L.POL_KIND_FLUID, dtype=np.uint8)
    kind[solid] = L.POL_KIND_SOLID
    kind[inlet] = L.POL_KIND_INLET
    kind[outlet] = L.POL_KIND_OUTLET

    flags, rev = link_flags(kind, periodic=(False, False, False))
    flags |= np.where(paper, np.uint32(L.POL_FLAG_POROUS | L.POL_FLAG_NO_LES), np.uint32(0))
    return Geometry(p, kind, paper, flags, rev, cone_top)


def _source_index(shape, e, periodic):
    """Per-axis source indices of a pull along e (x - e): periodic axes wrap,
    non-periodic axes clamp onto the cell's own plane.  Returns (iz, iy, ix,
    outside) with `outside` true where some non-periodic axis left the box."""
    nz, ny, nx = shape
    idx = []
    outs = []
    for n, ei, per in ((nz, e[2], periodic[2]), (ny, e[1], periodic[1]), (nx, e[0], periodic[0])):
        s = np.arange(n) - ei
        out = (s < 0) | (s >= n)
        s = np.mod(s, n) if per else np.clip(s, 0, n - 1)
        idx.append(s)
        outs.append(out & (not per))
    iz, iy, ix = idx
    oz, oy, ox = outs
    outside = oz[:, None, None] | oy[None, :, None] | ox[None, None, :]
    return iz, iy, ix, outside


def _shift_src(a: np.ndarray, e, periodic):
    """out[x] = a[source(x - e)] with wrap / clamp (see _source_index)."""
    iz, iy, ix, outside = _source_index(a.shape, e, periodic)
    return a[np.ix_(iz, iy, ix)], outside


def link_flags(kind: np.ndarray, periodic=(False, False, False)):
    """Bounce-back bits, kind bits and the adjoint reverse-gather word.

    A link q of a non-solid cell bounces when its source x - e_q lies outside
    a non-periodic box edge (except for inlet/outlet cells, whose out-of-box
    sources clamp onto their own plane) or when its (wrapped / clamped) source
    is solid.  No gather therefore ever reads a solid cell.
    """
    solid = kind == L.POL_KIND_SOLID
    inlet = kind == L.POL_KIND_INLET
    is_open = (kind == L.POL_KIND_INLET) | (kind == L.POL_KIND_OUTLET)
    flags = (kind.astype(np.uint32) << np.uint32(L.POL_KIND_SHIFT))
    reader_ok = (~solid) & (~inlet)     # cells that gather
    rev = np.zeros(kind.shape, dtype=np.uint32)
    for q in range(Q):
        e = tuple(int(v) for v in E[q])
        if q:
            src_solid, outside = _shift_src(solid, e, periodic)
            bounce = (src_solid | (outside & ~is_open)) & ~solid
            flags |= np.where(bounce, np.uint32(1 << q), np.uint32(0))
        # the normal reader of population q of cell y is y + e_q (pull along -e_q)
        rd, outside = _shift_src(reader_ok, tuple(-v for v in e), periodic)
        rd = rd & ~outside
        rev |= np.where(rd & ~solid, np.uint32(1 << q), np.uint32(0))
    # clamped readers exist only on non-periodic faces next to open cells
    n_z, n_y, n_x = kind.shape
    face = np.zeros(kind.shape, dtype=bool)
    for axis, per in zip((2, 1, 0), periodic):
        if per:
            continue
        sl0 = [slice(None)] * 3
        sl1 = [slice(None)] * 3
        sl0[axis] = 0
        sl1[axis] = -1
        face[tuple(sl0)] = True
        face[tuple(sl1)] = True
    rev |= np.where(face & ~solid, np.uint32(L.POL_RFLAG_CLAMP), np.uint32(0))
    return flags, rev


def periodic_box(n: tuple[int, int, int]):
    """Flags for a fully periodic all-fluid box (nx, ny, nz)."""
    nx, ny, nz = n
    kind = np.zeros((nz, ny, nx), dtype=np.uint8)
    return link_flags(kind, periodic=(True, True, True))


def params_dict(p: V60Params) -> dict:
    return asdict(p)
