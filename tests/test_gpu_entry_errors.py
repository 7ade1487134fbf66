"""Error behaviour of the entry points of the rows either side of the step:
invalid arguments and calls in the wrong state come back as POL_ERR_INVALID
/ POL_ERR_STATE with a message (pol_last_error) and launch nothing."""
from __future__ import annotations

import ctypes as C

import pytest
import torch

from pourover_amd import _lib as L
from pourover_amd.engine import HipContext, SolverConfig

pytestmark = pytest.mark.gpu

NX, NY, NZ = 12, 10, 8


def _ctx(force=False):
    return HipContext(SolverConfig(NX, NY, NZ, force_field=force))


def _status(ctx, fn, *args):
    st = fn(ctx.handle, *args)
    msg = (ctx.lib.pol_last_error(ctx.handle) or b"").decode()
    return st, msg


def _f32(*shape):
    return torch.zeros(shape, dtype=torch.float32, device="cuda")


def test_thermal_entry_errors():
    print("case: thermal entry errors", flush=True)
    ctx = _ctx()
    lib = ctx.lib
    s = L.stream_handle()
    st, _ = _status(ctx, lib.pol_thermal_step, 1, None, s)
    assert st == L.POL_ERR_STATE                                   # nothing bound
    g = [_f32(7, NZ, NY, NX) for _ in range(2)]
    T = [_f32(NZ, NY, NX) for _ in range(2)]
    hs = _f32(NZ, NY, NX)
    b = L.PolThermalBuffers()
    b.g[0], b.g[1] = g[0].data_ptr(), g[1].data_ptr()
    b.temperature[0], b.temperature[1] = T[0].data_ptr(), T[1].data_ptr()
    b.heat_source = hs.data_ptr()
    d = L.PolThermalDesc()
    d.omega = 2.5
    st, msg = _status(ctx, lib.pol_thermal_bind, C.byref(d), C.byref(b))
    assert st == L.POL_ERR_INVALID and "omega" in msg
    d.omega, d.options = 1.2, L.POL_THERMAL_CONVECTION
    st, msg = _status(ctx, lib.pol_thermal_bind, C.byref(d), C.byref(b))
    assert st == L.POL_ERR_INVALID and "velocity" in msg
    d.options = L.POL_THERMAL_RESET_SOURCE
    st, msg = _status(ctx, lib.pol_thermal_bind, C.byref(d), C.byref(b))
    assert st == L.POL_ERR_INVALID and "heat_source_base" in msg
    d.options = 0
    assert lib.pol_thermal_bind(ctx.handle, C.byref(d), C.byref(b)) == L.POL_OK
    assert lib.pol_thermal_boundary(ctx.handle, 4, hs.data_ptr(), 1.0, 0.0, s) == L.POL_ERR_INVALID
    assert lib.pol_thermal_step(ctx.handle, -1, None, s) == L.POL_ERR_INVALID


def test_multiphase_entry_errors():
    print("case: multiphase entry errors", flush=True)
    ctx = _ctx(force=False)
    lib = ctx.lib
    s = L.stream_handle()
    phi = [_f32(NZ, NY, NX) for _ in range(2)]
    aux = [_f32(NZ, NY, NX) for _ in range(3)]
    b = L.PolMultiphaseBuffers()
    b.phi[0], b.phi[1] = phi[0].data_ptr(), phi[0].data_ptr()   # the same buffer twice
    b.phi_new, b.mu, b.phase = aux[0].data_ptr(), aux[1].data_ptr(), aux[2].data_ptr()
    d = L.PolMultiphaseDesc()
    d.sigma, d.mobility, d.interface_width, d.rho_water, d.rho_air, d.dt = 0.0036, 0.001, 2.0, 1.0, 0.00125, 1.0
    st, _ = _status(ctx, lib.pol_multiphase_step, s)
    assert st == L.POL_ERR_STATE
    st, msg = _status(ctx, lib.pol_multiphase_bind, C.byref(d), C.byref(b))
    assert st == L.POL_ERR_INVALID and "phi" in msg
    b.phi[1] = phi[1].data_ptr()
    d.dt = float("nan")
    assert lib.pol_multiphase_bind(ctx.handle, C.byref(d), C.byref(b)) == L.POL_ERR_INVALID
    d.dt = 1.0
    assert lib.pol_multiphase_bind(ctx.handle, C.byref(d), C.byref(b)) == L.POL_OK
    st, msg = _status(ctx, lib.pol_multiphase_interface, L.POL_MP_APPLY, None, s)
    assert st == L.POL_ERR_STATE and "force" in msg                # no force field in this context
    assert lib.pol_multiphase_interface(ctx.handle, 0x80, None, s) == L.POL_ERR_INVALID


def test_forcing_entry_errors():
    print("case: forcing entry errors", flush=True)
    ctx = _ctx(force=False)
    lib = ctx.lib
    s = L.stream_handle()
    d = L.PolPourDesc()
    d.center_x, d.center_y, d.radius, d.stream_height = 5.0, 5.0, 4.0, 4.0
    d.velocity, d.flow_rate, d.dt, d.height = 0.05, 1.0, 1.0, 6
    st, msg = _status(ctx, lib.pol_pour_force, C.byref(d), s)
    assert st == L.POL_ERR_STATE and "force" in msg
    d.center_x = float("inf")
    phi = _f32(NZ, NY, NX)
    st, msg = _status(ctx, lib.pol_pour_phase, C.byref(d), phi.data_ptr(), s)
    assert st == L.POL_ERR_INVALID and "non-finite" in msg
    pf = _f32(3, NZ, NY, NX)
    assert lib.pol_pressure_force(ctx.handle, pf.data_ptr(), 3, 0.12, 1 / 3, s) == L.POL_ERR_INVALID
    st, msg = _status(ctx, lib.pol_pressure_force, pf.data_ptr(), 1, 0.12, 1 / 3, s)
    assert st == L.POL_ERR_STATE and "force" in msg
    assert lib.pol_pressure_force(ctx.handle, pf.data_ptr(), 0, 0.12, 1 / 3, s) == L.POL_OK  # store only
    st, msg = _status(ctx, lib.pol_pressure_force, pf.data_ptr(), L.POL_PRESSURE_SURFACE_TENSION, 0.12, 1 / 3, s)
    assert st == L.POL_ERR_INVALID and "modes 1 and 2" in msg     # nothing to add the tension to
    assert lib.pol_pressure_force(ctx.handle, pf.data_ptr(), 8, 0.12, 1 / 3, s) == L.POL_ERR_INVALID
    fctx = _ctx(force=True)
    st, msg = _status(fctx, lib.pol_pressure_force, pf.data_ptr(), 1 | L.POL_PRESSURE_SURFACE_TENSION, 0.12, 1 / 3, s)
    assert st == L.POL_ERR_STATE and "phase field" in msg          # no phase field bound
    cells = torch.zeros(4, dtype=torch.int32, device="cuda")
    st, msg = _status(ctx, lib.pol_filter_blockage, cells.data_ptr(), NX * NY * NZ + 1, phi.data_ptr(),
                      phi.data_ptr(), s)
    assert st == L.POL_ERR_INVALID and "paper cells" in msg
    assert lib.pol_filter_block_particles(ctx.handle, cells.data_ptr(), phi.data_ptr(), None, None, None, 0, 0.0, 0,
                                          s) == L.POL_ERR_INVALID          # scale must be positive
    assert lib.pol_density_drive(ctx.handle, phi.data_ptr(), 0.025, 0, 0.001, 2.0, 0.5, s) == L.POL_ERR_INVALID
    torch.cuda.synchronize()
