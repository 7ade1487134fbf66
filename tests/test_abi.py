"""The C ABI without a GPU: the library loads, exports every function that
include/pourover_hip.h declares, and the ctypes mirror (_lib.py) agrees with
the header on constants and struct layouts (offsets from a gcc-compiled probe
of the header itself).  No compute entry point is called."""
from __future__ import annotations

import ctypes as C
import re
import shutil
import subprocess
from pathlib import Path

import pytest

from pourover_amd import _lib as L

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "pourover_hip.h"


def _header_text() -> str:
    txt = HEADER.read_text()
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    return txt


def _declared_functions() -> list[str]:
    return sorted(set(re.findall(r"\b(pol_[a-z0-9_]+)\s*\(", _header_text())))


def _defines() -> dict[str, int]:
    out = {}
    for name, val in re.findall(r"#define\s+(POL_[A-Z0-9_]+)\s+(0x[0-9A-Fa-f]+|\d+)u?\b", _header_text()):
        out[name] = int(val, 0)
    return out


def _enums() -> dict[str, int]:
    out = {}
    for body in re.findall(r"enum\s*\{(.*?)\}", _header_text(), flags=re.S):
        nxt = 0
        for item in body.split(","):
            item = item.strip()
            if not item:
                continue
            if "=" in item:
                k, v = (s.strip() for s in item.split("="))
                nxt = int(v, 0)
            else:
                k = item
            out[k] = nxt
            nxt += 1
    return out


def test_library_loads_and_exports_every_declared_function():
    lib = L.load()
    decl = _declared_functions()
    assert len(decl) >= 20
    missing = [n for n in decl if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.pol_version() == L.POL_ABI_VERSION


def test_ctypes_signatures_cover_the_header():
    assert sorted(L.SIGNATURES) == _declared_functions()


def test_status_strings_need_no_device():
    lib = L.load()
    for st in range(7):
        assert lib.pol_status_string(st)
    # argument validation happens before any HIP call
    assert lib.pol_create(None, None) == L.POL_ERR_INVALID


def test_constants_match_header():
    d = _defines()
    e = _enums()
    for name, val in list(d.items()) + list(e.items()):
        if name in ("POUROVER_HIP_H",):
            continue
        if name.startswith("POL_RED_"):
            continue  # slots: checked below against RED_SLOTS
        assert hasattr(L, name), name
        assert getattr(L, name) == val, (name, getattr(L, name), val)
    red = sorted((v, k) for k, v in d.items() if k.startswith("POL_RED_"))
    assert [v for v, _ in red] == list(range(L.POL_NRED)) == list(range(len(L.RED_SLOTS)))


_PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "pourover_hip.h"
#define F(T, m) printf(#T " " #m " %zu %zu\n", offsetof(T, m), sizeof(((T *)0)->m));
#define S(T) printf(#T " __size__ %zu %zu\n", sizeof(T), (size_t)0);
int main(void) {
  S(pol_desc) F(pol_desc, nx) F(pol_desc, ny) F(pol_desc, nz) F(pol_desc, storage) F(pol_desc, options)
  F(pol_desc, device) F(pol_desc, tau) F(pol_desc, cs_smag) F(pol_desc, nu_sgs_max) F(pol_desc, gravity)
  F(pol_desc, darcy_K) F(pol_desc, forch_CF) F(pol_desc, inlet_rho) F(pol_desc, inlet_u) F(pol_desc, outlet_rho)
  S(pol_buffers) F(pol_buffers, f) F(pol_buffers, cell_flags) F(pol_buffers, force) F(pol_buffers, perm)
  F(pol_buffers, rho) F(pol_buffers, u) F(pol_buffers, active_chunks) F(pol_buffers, n_chunks)
  F(pol_buffers, chunk_cells) F(pol_buffers, particle_drag) F(pol_buffers, chunk_records)
  F(pol_buffers, stored_cells) F(pol_buffers, layout_flags) F(pol_buffers, nu_sgs) F(pol_buffers, les_phase)
  S(pol_particle_desc) F(pol_particle_desc, count) F(pol_particle_desc, drag_law) F(pol_particle_desc, rho_fluid)
  F(pol_particle_desc, nu_fluid) F(pol_particle_desc, drag_max_per_mass) F(pol_particle_desc, relax_alpha)
  F(pol_particle_desc, gravity) F(pol_particle_desc, dt) F(pol_particle_desc, max_disp)
  F(pol_particle_desc, wall_damping) F(pol_particle_desc, reaction) F(pol_particle_desc, reserved)
  S(pol_particle_buffers) F(pol_particle_buffers, pos) F(pol_particle_buffers, vel) F(pol_particle_buffers, radius)
  F(pol_particle_buffers, mass) F(pol_particle_buffers, active) F(pol_particle_buffers, drag_new)
  F(pol_particle_buffers, drag_old) F(pol_particle_buffers, drag) F(pol_particle_buffers, u_fluid)
  F(pol_particle_buffers, re_p) F(pol_particle_buffers, cd) F(pol_particle_buffers, drag_field)
  F(pol_particle_buffers, u_field) F(pol_particle_buffers, reaction_field)
  S(pol_adjoint_buffers) F(pol_adjoint_buffers, lam) F(pol_adjoint_buffers, grad_perm)
  F(pol_adjoint_buffers, rev_flags)
  S(pol_stats_opts) F(pol_stats_opts, prev_rho) F(pol_stats_opts, prev_u) F(pol_stats_opts, phase)
  F(pol_stats_opts, z_top) F(pol_stats_opts, z_mid) F(pol_stats_opts, z_outlet) F(pol_stats_opts, with_populations)
  S(pol_thermal_desc) F(pol_thermal_desc, omega) F(pol_thermal_desc, options) F(pol_thermal_desc, reserved)
  S(pol_thermal_buffers) F(pol_thermal_buffers, g) F(pol_thermal_buffers, temperature)
  F(pol_thermal_buffers, heat_source) F(pol_thermal_buffers, velocity) F(pol_thermal_buffers, conductivity)
  F(pol_thermal_buffers, heat_source_base)
  S(pol_multiphase_desc) F(pol_multiphase_desc, sigma) F(pol_multiphase_desc, mobility)
  F(pol_multiphase_desc, interface_width) F(pol_multiphase_desc, rho_water) F(pol_multiphase_desc, rho_air)
  F(pol_multiphase_desc, dt)
  S(pol_multiphase_buffers) F(pol_multiphase_buffers, phi) F(pol_multiphase_buffers, phi_new)
  F(pol_multiphase_buffers, mu) F(pol_multiphase_buffers, phase)
  S(pol_interface_fields) F(pol_interface_fields, grad_phi) F(pol_interface_fields, grad_mu)
  F(pol_interface_fields, normal) F(pol_interface_fields, curvature) F(pol_interface_fields, surface_force)
  S(pol_pour_desc) F(pol_pour_desc, center_x) F(pol_pour_desc, center_y) F(pol_pour_desc, radius)
  F(pol_pour_desc, stream_height) F(pol_pour_desc, velocity) F(pol_pour_desc, flow_rate) F(pol_pour_desc, dt)
  F(pol_pour_desc, height) F(pol_pour_desc, reserved)
  return 0;
}
"""

_CTYPES = {"pol_desc": L.PolDesc, "pol_buffers": L.PolBuffers, "pol_particle_desc": L.PolParticleDesc,
           "pol_particle_buffers": L.PolParticleBuffers, "pol_adjoint_buffers": L.PolAdjointBuffers,
           "pol_stats_opts": L.PolStatsOpts, "pol_thermal_desc": L.PolThermalDesc,
           "pol_thermal_buffers": L.PolThermalBuffers, "pol_multiphase_desc": L.PolMultiphaseDesc,
           "pol_multiphase_buffers": L.PolMultiphaseBuffers, "pol_interface_fields": L.PolInterfaceFields,
           "pol_pour_desc": L.PolPourDesc}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "probe.c"
    exe = tmp_path / "probe"
    src.write_text(_PROBE)
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, stdout=subprocess.PIPE, text=True).stdout.split("\n")
    seen = {k: set() for k in _CTYPES}
    for line in filter(None, out):
        tname, field, off, size = line.split()
        ct = _CTYPES[tname]
        if field == "__size__":
            assert C.sizeof(ct) == int(off), tname
            continue
        desc = getattr(ct, field)
        assert desc.offset == int(off), (tname, field, desc.offset, off)
        assert desc.size == int(size), (tname, field, desc.size, size)
        seen[tname].add(field)
    for tname, ct in _CTYPES.items():
        assert seen[tname] == {f[0] for f in ct._fields_}, tname


def test_missing_library_fails_loudly(tmp_path):
    """No CPU fallback: without the HIP library the product path raises."""
    import os
    import sys

    code = ("import sys; sys.path.insert(0, %r); import __graft_entry__ as g; g.load_package(); "
            "from pourover_amd import _lib\n"
            "try:\n    _lib.load()\nexcept ImportError as e:\n    print('raised', e)\n    sys.exit(0)\n"
            "sys.exit(3)") % str(ROOT)
    env = dict(os.environ, POUROVER_HIP_LIB=str(tmp_path / "absent.so"))
    p = subprocess.run([sys.executable, "-c", code], env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "raised" in p.stdout


def test_noalloc_reduction_rejects_before_any_device_call():
    """pol_reduce_noalloc checks its arguments on the host: null arguments and a
    state of more work items than the static scratch holds come back as
    POL_ERR_INVALID with a message, before any HIP call (none is made here:
    this runs without a device)."""
    lib = L.load()
    d = L.PolDesc()
    d.nx = d.ny = d.nz = 512  # 134 217 728 cells: 524 288 blocks of 256
    d.storage = L.POL_F32
    d.tau = 0.6
    b = L.PolBuffers()
    b.f[0] = 0x1000  # never dereferenced: the size check comes first
    out = C.c_void_p(0x2000)
    assert lib.pol_reduce_noalloc(None, C.byref(b), out, None) == L.POL_ERR_INVALID
    assert b"null argument" in lib.pol_last_error(None)
    st = lib.pol_reduce_noalloc(C.byref(d), C.byref(b), out, None)
    assert st == L.POL_ERR_INVALID
    msg = lib.pol_last_error(None).decode()
    assert "exceed the static scratch" in msg and str(L.POL_NOALLOC_MAX_BLOCKS * 256) in msg
    d.tau = 0.4  # invalid physics
    assert lib.pol_reduce_noalloc(C.byref(d), C.byref(b), out, None) == L.POL_ERR_INVALID
    assert b"invalid descriptor" in lib.pol_last_error(None)
