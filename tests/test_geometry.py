"""V60 geometry and flag words (CPU): integer results pinned by invariants and
committed digests (tests/golden/golden.json, written by tests/golden/make_golden.py)."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest
from scipy import ndimage

from pourover_amd import _lib as L
from pourover_amd.geometry import V60Params, build_v60, link_flags
from pourover_amd.lattice import E, OPP, Q

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "golden.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as fh:
        return json.load(fh)


@pytest.fixture(scope="module")
def g48():
    return build_v60(V60Params.scaled(48))


def _kind(flags):
    return (flags >> L.POL_KIND_SHIFT) & L.POL_KIND_MASK


@pytest.mark.parametrize("n", [32, 48, 128])
def test_digest_matches_golden(golden, n):
    g = build_v60(V60Params() if n == 224 else V60Params.scaled(n))
    ref = golden["geometry"][str(n)]
    assert g.counts() == ref["counts"]
    assert g.digest() == ref["sha256"]


def test_digest_224(golden):
    g = build_v60()
    ref = golden["geometry"]["224"]
    assert g.counts() == ref["counts"]
    assert g.digest() == ref["sha256"]
    # the 224^3 V60 cup: outlet hole radius 8 -> outlet plane = discrete disk r <= 8 around (111.5, 111.5)
    yy, xx = np.mgrid[0:224, 0:224]
    disk = ((xx - 111.5) ** 2 + (yy - 111.5) ** 2) <= 64.0
    assert g.counts()["outlet"] == int(disk.sum())
    jet = ((xx - 111.5) ** 2 + (yy - 111.5) ** 2) <= 16.0
    assert g.counts()["inlet"] == int(jet.sum())


def _check_no_solid_gather(flags, periodic):
    kind = _kind(flags)
    nz, ny, nx = kind.shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    gathers = (kind != L.POL_KIND_SOLID) & (kind != L.POL_KIND_INLET)
    dims = (nx, ny, nz)
    for q in range(1, Q):
        bounce = (flags >> np.uint32(q)) & 1
        src = []
        for c, e, n, per in ((x, E[q, 0], nx, periodic[0]), (y, E[q, 1], ny, periodic[1]), (z, E[q, 2], nz, periodic[2])):
            s = c - e
            s = np.mod(s, n) if per else np.clip(s, 0, n - 1)
            src.append(s)
        src_kind = kind[src[2], src[1], src[0]]
        bad = gathers & (bounce == 0) & (src_kind == L.POL_KIND_SOLID)
        assert not bad.any(), "link %d gathers a solid cell at %s" % (q, np.argwhere(bad)[:3])


def test_no_gather_reads_a_solid_cell(g48):
    _check_no_solid_gather(g48.flags, (False, False, False))


def test_bounce_bits_only_on_open_cells_and_mirror_walls(g48):
    kind = _kind(g48.flags)
    solid = kind == L.POL_KIND_SOLID
    assert ((g48.flags[solid] & L.POL_LINK_MASK) == 0).all()
    assert ((g48.flags & 1) == 0).all()          # rest population never bounces


def test_periodic_box_has_no_walls():
    flags, rev = link_flags(np.zeros((6, 5, 4), dtype=np.uint8), periodic=(True, True, True))
    assert (flags == 0).all()
    assert (rev == (1 << Q) - 1).all()           # every population has its normal reader


def test_outlet_and_inlet_planes(g48):
    kind = _kind(g48.flags)
    assert (kind[0][kind[0] != L.POL_KIND_SOLID] == L.POL_KIND_OUTLET).all()
    assert (kind[-1][kind[-1] != L.POL_KIND_SOLID] == L.POL_KIND_INLET).all()
    assert (kind[1:-1] != L.POL_KIND_OUTLET).all() and (kind[1:-1] != L.POL_KIND_INLET).all()


@pytest.mark.parametrize("n", [48, 224])
def test_filter_separates_inlet_from_outlet(n):
    """Every D3Q19 path from the jet to the outlet crosses the porous paper:
    the non-porous fluid splits (18-connectivity = the D3Q19 links) into a
    component holding the inlet and one holding the outlet."""
    g = build_v60(V60Params() if n == 224 else V60Params.scaled(n))
    kind = g.kind
    open_fluid = (kind != L.POL_KIND_SOLID) & ~g.porous
    lab, _ = ndimage.label(open_fluid, structure=ndimage.generate_binary_structure(3, 2))
    lin = set(np.unique(lab[kind == L.POL_KIND_INLET])) - {0}
    lout = set(np.unique(lab[kind == L.POL_KIND_OUTLET])) - {0}
    assert lin and lout and not (lin & lout)
    # ... while the whole non-solid region (paper included) is connected
    lab2, _ = ndimage.label(kind != L.POL_KIND_SOLID, structure=ndimage.generate_binary_structure(3, 2))
    a = set(np.unique(lab2[kind == L.POL_KIND_INLET]))
    b = set(np.unique(lab2[kind == L.POL_KIND_OUTLET]))
    assert a == b


def test_reverse_flags_are_the_transpose_of_gathers(g48):
    """rev bit p of y set  <=>  cell y + e_p gathers population p of y by a normal link."""
    flags, rev = g48.flags, g48.rev_flags
    kind = _kind(flags)
    nz, ny, nx = kind.shape
    for p in range(Q):
        ex, ey, ez = E[p]
        zs, ys, xs = np.nonzero((rev >> np.uint32(p)) & 1)
        rz, ry, rx = zs + ez, ys + ey, xs + ex
        assert ((rz >= 0) & (rz < nz) & (ry >= 0) & (ry < ny) & (rx >= 0) & (rx < nx)).all()
        rk = kind[rz, ry, rx]
        assert ((rk != L.POL_KIND_SOLID) & (rk != L.POL_KIND_INLET)).all()
        assert (((flags[rz, ry, rx] >> np.uint32(p)) & 1) == 0).all()
        assert (kind[zs, ys, xs] != L.POL_KIND_SOLID).all()


def test_normal_reader_and_bounce_are_exclusive(g48):
    """For every non-solid cell x and direction p, f_p(x) is read either by
    x + e_p (rev bit p) or by x through its bounce link opp(p), never both:
    the adjoint's reverse gather does one load per direction on that basis."""
    kind = g48.kind
    fluid = kind != L.POL_KIND_SOLID
    for p in range(1, 19):
        rev_p = (g48.rev_flags >> np.uint32(p)) & np.uint32(1)
        bounce = (g48.flags >> np.uint32(OPP[p])) & np.uint32(1)
        assert not np.any(fluid & (rev_p == 1) & (bounce == 1)), p


def test_porous_cells_have_les_off(g48):
    por = (g48.flags & L.POL_FLAG_POROUS) != 0
    assert por.any()
    assert ((g48.flags[por] & L.POL_FLAG_NO_LES) != 0).all()
    assert (_kind(g48.flags)[por] == L.POL_KIND_FLUID).all()


@pytest.fixture(scope="module")
def g224():
    return build_v60()


def test_cone_half_angle_is_30_degrees(g224):
    """The cup wall measured from the 224^3 mask: in every plane between the
    cone base and the top of the cone the open radius (largest centre
    distance of a non-solid cell) grows linearly with z; the least-squares
    slope is tan(half-angle), and the half-angle is 30 deg (the V60 workload:
    30 deg cone, outlet hole radius 8 cells)."""
    n = 224
    kind = g224.kind
    zb = g224.params.hole_length
    yy, xx = np.mgrid[0:n, 0:n]
    r = np.sqrt((xx - 111.5) ** 2 + (yy - 111.5) ** 2)
    zs = np.arange(zb + 4, g224.cone_top_z - 4)
    radii = np.array([r[kind[z] != L.POL_KIND_SOLID].max() for z in zs])
    slope, icpt = np.polyfit(zs, radii, 1)
    angle = np.degrees(np.arctan(slope))
    assert abs(angle - 30.0) < 0.25, angle
    # a cone, not a staircase or a curve: within one cell of the line everywhere
    assert np.abs(radii - (slope * zs + icpt)).max() < 1.0
    # and the line meets the outlet hole (radius 8) at the cone base
    assert abs(slope * zb + icpt - 8.0) < 1.0
    # above the cone top the wall is vertical (the cup's rim radius)
    above = [r[kind[z] != L.POL_KIND_SOLID].max() for z in range(g224.cone_top_z + 1, n - 2)]
    assert max(above) - min(above) < 1e-9


def test_filter_layer_is_two_cells_thick(g224):
    """The porous filter layer measured from the 224^3 mask: along radial
    lines at several heights and along the axis through the base disk, each
    crossing of the paper is two porous cells (the 2-cell filter layer)."""
    n = 224
    por = g224.porous
    zb = g224.params.hole_length

    def runs(line):
        """lengths of the runs of True in a 1-D boolean array"""
        d = np.diff(np.concatenate(([0], line.astype(np.int8), [0])))
        return list(np.flatnonzero(d == -1) - np.flatnonzero(d == 1))

    # radial lines along +x, -x, +y, -y through the axis (row 111: 0.5 cells off it)
    for z in (zb + 20, zb + 60, zb + 100, zb + 140):
        for line in (por[z, 111, 112:], por[z, 111, :112][::-1], por[z, 112:, 111], por[z, :112, 111][::-1]):
            assert runs(line) == [2], (z, runs(line))
    # the base disk closes the cone: two cells thick along z at the axis
    for (y, x) in ((111, 111), (112, 112), (109, 113)):  # inside the disk (r < 5 at the base)
        col = por[:, y, x]
        assert runs(col) == [2], runs(col)
        assert col[zb] and col[zb + 1]
    # the layer is porous and has LES off; nothing else is
    assert np.array_equal(por, (g224.flags & L.POL_FLAG_POROUS) != 0)
