"""Pins of the oracle's open-boundary rules (CPU, FP64).

The oracle (oracle/lbm_oracle.c) gives open cells three rules:
  * an inlet cell's post-collision state is feq(inlet_rho, inlet_u);
  * an outlet cell's is feq(outlet_rho, u), u the half-force velocity of the
    populations it gathered;
  * an out-of-box source of an inlet/outlet cell on a non-periodic face is
    clamped onto the cell's own plane on that axis (zero gradient), the other
    axes shifting as usual.
Each check below is fixed by a closed form or a conservation law, not by the
oracle's own formula.  The reference's boundary kernels are
src/physics/boundary_conditions.py:186-324 (inlet/outlet equilibria and the
extrapolated open faces).
"""
from __future__ import annotations

import numpy as np

from pourover_amd import _lib as L
from pourover_amd.geometry import link_flags
from pourover_amd.lattice import E, equilibrium

PER_XY = L.POL_PERIODIC_X | L.POL_PERIODIC_Y


def _uniform(shape, rho0, u0):
    nz, ny, nx = shape
    rho = np.full(shape, rho0)
    u = np.broadcast_to(np.asarray(u0, dtype=np.float64)[:, None, None, None], (3,) + shape).copy()
    return rho, u


def test_inlet_and_outlet_rules_closed_form(oracle):
    """From a uniform equilibrium feq(rho0, u0) under a uniform force F every
    gathered state (clamped sources included) is feq(rho0, u0), whose half-force
    velocity is u0 + F/(2 rho0).  One step then gives inlet cells
    feq(inlet_rho, inlet_u) and outlet cells feq(outlet_rho, u0 + F/(2 rho0))."""
    nx, ny, nz = 4, 5, 6
    kind = np.zeros((nz, ny, nx), np.uint8)
    kind[-1] = L.POL_KIND_INLET
    kind[0] = L.POL_KIND_OUTLET
    flags, _ = link_flags(kind, periodic=(True, True, False))
    rho0, u0 = 1.01, np.array([0.004, -0.003, -0.02])
    F = np.array([1e-5, 2e-5, -3e-5])
    rin, uin, rout = 0.995, np.array([0.001, 0.002, -0.03]), 0.998
    d = oracle.make_desc(nx, ny, nz, tau=0.7, options=PER_XY | L.POL_USE_FLAGS | L.POL_GRAVITY, gravity=tuple(F),
                         inlet_rho=rin, inlet_u=tuple(uin), outlet_rho=rout)
    rho, u = _uniform((nz, ny, nx), rho0, u0)
    f0 = oracle.init_equilibrium(nx, ny, nz, rho, u)
    f1, rho1, u1, red = oracle.step(d, f0, flags, want_macro=True)
    feq_in = equilibrium(np.array([rin]), uin.reshape(3, 1))[:, 0]
    uh = u0 + F / (2 * rho0)
    feq_out = equilibrium(np.array([rout]), uh.reshape(3, 1))[:, 0]
    for y in range(ny):
        for x in range(nx):
            assert np.allclose(f1[:, -1, y, x], feq_in, rtol=1e-14, atol=0)
            assert np.allclose(f1[:, 0, y, x], feq_out, rtol=1e-13, atol=0)
    # reported moments: the imposed inlet state; the gathered outlet state
    assert np.allclose(rho1[-1], rin, rtol=1e-15) and np.allclose(u1[:, -1], uin[:, None, None], rtol=1e-15)
    assert np.allclose(rho1[0], rho0, rtol=1e-14)
    assert np.allclose(u1[:, 0], uh[:, None, None], rtol=1e-12)
    # inlet flux slot: sum over inlet cells of -rho u_z
    assert abs(red[L.RED_SLOTS.index("inlet_flux")] - nx * ny * rin * 0.03) < 1e-14


def test_uniform_duct_stays_uniform_influx_equals_outflux(oracle):
    """A uniform stream feq(rho0, u0) through a duct (periodic x, y; inlet plane
    on top imposing (rho0, u0); outlet plane at the bottom with outlet_rho =
    rho0) is a fixed point: every cell keeps feq(rho0, u0), and the inlet and
    outlet fluxes are equal, area * rho0 * |u_z|."""
    nx, ny, nz = 5, 4, 9
    rho0, u0 = 1.0, np.array([0.0, 0.0, -0.03])
    kind = np.zeros((nz, ny, nx), np.uint8)
    kind[-1] = L.POL_KIND_INLET
    kind[0] = L.POL_KIND_OUTLET
    flags, _ = link_flags(kind, periodic=(True, True, False))
    d = oracle.make_desc(nx, ny, nz, tau=0.6, options=PER_XY | L.POL_USE_FLAGS, inlet_rho=rho0, inlet_u=tuple(u0),
                         outlet_rho=rho0)
    rho, u = _uniform((nz, ny, nx), rho0, u0)
    f = oracle.init_equilibrium(nx, ny, nz, rho, u)
    f0 = f.copy()
    f, _, _, _ = oracle.run(d, f, 200, flags=flags)
    assert np.allclose(f, f0, rtol=1e-13, atol=0)  # FP64 round-off only
    _, _, red = oracle.moments(d, f, flags)
    fin, fout = red[L.RED_SLOTS.index("inlet_flux")], red[L.RED_SLOTS.index("outlet_flux")]
    assert abs(fin - nx * ny * rho0 * 0.03) < 1e-14
    assert abs(fout - fin) <= 1e-12 * fin


def test_open_faces_keep_a_uniform_flow(oracle):
    """Outlet cells on every face of a non-periodic box (edges and corners
    included), outlet_rho = rho0: a uniform 3-D flow stays exactly uniform under
    the zero-gradient clamp, and the mass is unchanged."""
    n = 6
    rho0, u0 = 1.02, np.array([0.011, -0.017, 0.013])
    kind = np.full((n, n, n), L.POL_KIND_OUTLET, np.uint8)
    kind[1:-1, 1:-1, 1:-1] = L.POL_KIND_FLUID
    flags, _ = link_flags(kind, periodic=(False, False, False))
    assert (flags & np.uint32(L.POL_LINK_MASK)).max() == 0          # no link bounces
    d = oracle.make_desc(n, n, n, tau=0.65, options=L.POL_USE_FLAGS, outlet_rho=rho0)
    rho, u = _uniform((n, n, n), rho0, u0)
    f = oracle.init_equilibrium(n, n, n, rho, u)
    f0 = f.copy()
    f, _, _, _ = oracle.run(d, f, 100, flags=flags)
    assert np.allclose(f, f0, rtol=1e-13, atol=0)  # FP64 round-off only
    _, _, red = oracle.moments(d, f, flags)
    assert abs(red[0] - n**3 * rho0) < 1e-12


def test_clamp_keeps_the_other_axes_shift(oracle):
    """A state that does not vary along x: clamping an out-of-box x source onto
    the face plane reads the value a periodic wrap would read, provided the y/z
    parts of the source still shift by -e.  So outlet planes at x = 0 and
    x = nx - 1 gather exactly what a periodic-x box gathers, for an arbitrary
    (non-equilibrium) y/z-dependent state."""
    nx, ny, nz = 5, 7, 6
    rng = np.random.default_rng(11)
    col = 1.0 / 19 + 0.01 * rng.standard_normal((19, nz, ny, 1))
    f = np.ascontiguousarray(np.broadcast_to(col, (19, nz, ny, nx)))
    kind = np.zeros((nz, ny, nx), np.uint8)
    kind[:, :, 0] = kind[:, :, -1] = L.POL_KIND_OUTLET
    fl_open, _ = link_flags(kind, periodic=(False, True, True))
    fl_per, _ = link_flags(kind, periodic=(True, True, True))
    assert (fl_open & np.uint32(L.POL_LINK_MASK)).max() == 0
    d_open = oracle.make_desc(nx, ny, nz, tau=0.6, options=L.POL_PERIODIC_Y | L.POL_PERIODIC_Z | L.POL_USE_FLAGS)
    d_per = oracle.make_desc(nx, ny, nz, tau=0.6, options=PER_XY | L.POL_PERIODIC_Z | L.POL_USE_FLAGS)
    r_o, u_o, _ = oracle.moments(d_open, f, fl_open)
    r_p, u_p, _ = oracle.moments(d_per, f, fl_per)
    assert np.array_equal(r_o, r_p) and np.array_equal(u_o, u_p)
    # the state does depend on y and z, so a clamp that also froze y/z would differ
    frozen = f.sum(axis=0)
    assert np.abs(r_o[:, :, 0] - frozen[:, :, 0]).max() > 1e-3


def test_steady_open_flow_carries_one_flux_through_every_layer(oracle):
    """Steady flow from an inlet patch in a lid on top to an outlet plane at the
    bottom (periodic x, y).  Without a body force BGK keeps each cell's
    momentum, so at a fixed point the z momentum of every layer equals the
    population flux through its faces; the flux is the same through every
    plane (steady continuity), hence every layer, the outlet layer with its
    clamped sources included, carries the same sum of rho u_z.  A clamp onto the
    wrong plane, a lost population or an inlet with the wrong sign breaks it."""
    nx, ny, nz = 6, 6, 8
    kind = np.zeros((nz, ny, nx), np.uint8)
    kind[-1] = L.POL_KIND_SOLID
    kind[-1, 2:4, 2:4] = L.POL_KIND_INLET
    kind[0] = L.POL_KIND_OUTLET
    flags, _ = link_flags(kind, periodic=(True, True, False))
    d = oracle.make_desc(nx, ny, nz, tau=0.8, options=PER_XY | L.POL_USE_FLAGS, inlet_rho=1.0,
                         inlet_u=(0.0, 0.0, -0.02), outlet_rho=0.999)
    f = oracle.init_equilibrium(nx, ny, nz)
    for _ in range(200):
        prev = f
        f, _, _, _ = oracle.run(d, f, 200, flags=flags)
        if np.abs(f - prev).max() < 2e-15:      # converged to FP64 round-off
            break
    assert np.abs(f - prev).max() < 2e-15
    rho, u, red = oracle.moments(d, f, flags)
    layer = (rho * u[2]).sum(axis=(1, 2))[:-1]            # layers 0 .. nz-2 (below the lid)
    phi = layer[0]
    assert phi < -1e-3                                     # downward flow
    assert np.abs(layer - phi).max() <= 1e-11 * abs(phi)
    # the flow is three-dimensional: the outlet profile is not uniform
    assert np.ptp(u[2, 0]) > 1e-4 * abs(u[2, 0]).max()
    # and the outlet flux slot is that layer flux
    assert abs(red[L.RED_SLOTS.index("outlet_flux")] + phi) <= 1e-11 * abs(phi)


def test_zero_gradient_face_gathers_linear_rest_state(oracle):
    """Rest state with a density linear along one axis, rho(i) = 1 + a i, open
    (outlet) planes at both ends of that axis, periodic along the others.  The
    population entering a face cell from outside equals the one the face plane
    itself sends in that direction, so with the D3Q19 weights (2/3 of the mass
    in links along neither way of the axis, 1/6 each way) a face cell gathers
    rho = rho_face +- a/6 and rho u = -a/6 along the axis (interior cells:
    rho(i) and -a/3).  Clamping onto the next plane inward, or dropping the
    entering populations, changes both."""
    a = 2e-3
    for axis in range(3):                       # 0: x, 1: y, 2: z
        dims = [5, 6, 7]
        n = dims[axis]
        nx, ny, nz = dims
        shape = (nz, ny, nx)
        kind = np.zeros(shape, np.uint8)
        sl_lo = [slice(None)] * 3
        sl_hi = [slice(None)] * 3
        sl_lo[2 - axis] = 0
        sl_hi[2 - axis] = -1
        kind[tuple(sl_lo)] = L.POL_KIND_OUTLET
        kind[tuple(sl_hi)] = L.POL_KIND_OUTLET
        periodic = [True, True, True]
        periodic[axis] = False
        flags, _ = link_flags(kind, periodic=tuple(periodic))
        opts = L.POL_USE_FLAGS
        for k, bit in enumerate((L.POL_PERIODIC_X, L.POL_PERIODIC_Y, L.POL_PERIODIC_Z)):
            if periodic[k]:
                opts |= bit
        d = oracle.make_desc(nx, ny, nz, tau=0.6, options=opts)
        i = np.arange(n, dtype=np.float64)
        prof = 1.0 + a * i
        rho = np.broadcast_to(prof.reshape([n if 2 - ax == 2 - axis else 1 for ax in (2, 1, 0)]), shape).copy()
        f = oracle.init_equilibrium(nx, ny, nz, rho, np.zeros((3,) + shape))
        rg, ug, _ = oracle.moments(d, f, flags)
        rg_line = np.moveaxis(rg, 2 - axis, 0).reshape(n, -1)
        j_line = np.moveaxis(rg * ug[axis], 2 - axis, 0).reshape(n, -1)
        exp_r = prof.copy()
        exp_r[0] += a / 6
        exp_r[-1] -= a / 6
        exp_j = np.full(n, -a / 3)
        exp_j[0] = exp_j[-1] = -a / 6
        assert np.allclose(rg_line, exp_r[:, None], rtol=0, atol=1e-15), axis
        assert np.allclose(j_line, exp_j[:, None], rtol=0, atol=1e-16), axis
        others = [c for c in range(3) if c != axis]
        assert np.abs(ug[others]).max() < 1e-16
