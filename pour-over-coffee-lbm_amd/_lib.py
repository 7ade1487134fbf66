"""ctypes binding of libpourover_hip.so (the C ABI of include/pourover_hip.h).

The library is loaded after torch so that both share torch's HIP runtime
(libamdhip64.so.7 is matched by soname).  A missing library raises
ImportError: there is no CPU fallback for the hot path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import torch  # noqa: F401  (must precede the library: one HIP runtime per process)

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libpourover_hip.so"

# ---- constants (mirror include/pourover_hip.h) ------------------------------
POL_ABI_VERSION = 5
POL_Q = 19
POL_NRED = 12
POL_NOALLOC_MAX_BLOCKS = 65536
POL_CHUNK = 32  # cells per launch chunk
POL_RECORD_WORDS = 12  # uint32 per compact-layout row record
POL_PIECE_PLAIN = 0x80000000  # piece word bit: every cell of the piece has flag word 0
POL_LAYOUT_NO_X_FACE = 1  # pol_buffers.layout_flags: no run reaches x = 0 or nx - 1
POL_OK, POL_ERR_INVALID, POL_ERR_HIP, POL_ERR_ALLOC, POL_ERR_NO_DEVICE, POL_ERR_STATE, POL_ERR_UNSUPPORTED = range(7)
POL_F32, POL_F16, POL_BF16, POL_F64 = range(4)
POL_PERIODIC_X = 0x1
POL_PERIODIC_Y = 0x2
POL_PERIODIC_Z = 0x4
POL_USE_FLAGS = 0x8
POL_LES = 0x10
POL_FORCE_FIELD = 0x20
POL_PERM_FIELD = 0x40
POL_GRAVITY = 0x80
POL_PARTICLE_FIELD = 0x100
POL_LES_FD = 0x200
POL_LINK_MASK = 0x7FFFF
POL_KIND_SHIFT = 24
POL_KIND_MASK = 0x7
POL_KIND_FLUID, POL_KIND_SOLID, POL_KIND_INLET, POL_KIND_OUTLET = 0, 1, 2, 3
POL_FLAG_POROUS = 0x08000000
POL_FLAG_NO_LES = 0x10000000
POL_RFLAG_CLAMP = 0x80000000
POL_DRAG_TEXTBOOK = 0
POL_DRAG_REFERENCE = 1
POL_REACTION_IMPLICIT = 0
POL_REACTION_EXPLICIT = 1
POL_WANT_MACRO = 0x1
POL_WANT_REDUCE = 0x2
RED_SLOTS = ("mass", "mom_x", "mom_y", "mom_z", "kinetic_energy", "max_velocity", "nonfinite_cells",
             "outlet_flux", "cells", "rho_min", "rho_max", "inlet_flux")

# field statistics slots (pol_field_stats)
POL_STAT_CELLS = 0
POL_STAT_SUM_DRHO = 1
POL_STAT_SUM_DRHO2 = 2
POL_STAT_MAX_ABS_DRHO = 3
POL_STAT_MOM_X = 4
POL_STAT_MOM_Y = 5
POL_STAT_MOM_Z = 6
POL_STAT_SUM_UMAG = 7
POL_STAT_SUM_UMAG2 = 8
POL_STAT_MAX_UMAG = 9
POL_STAT_SUM_MACH_ACTIVE = 10
POL_STAT_ACTIVE = 11
POL_STAT_MACH_GT = 12
POL_STAT_SUM_UX = 13
POL_STAT_SUM_UY = 14
POL_STAT_SUM_UZ = 15
POL_STAT_SUM_UX2 = 16
POL_STAT_SUM_UY2 = 17
POL_STAT_SUM_UZ2 = 18
POL_STAT_SUM_UMAG_TOP = 19
POL_STAT_SUM_UMAG_MID = 20
POL_STAT_SUM_UMAG_BOT = 21
POL_STAT_OUTLET_DOWN = 22
POL_STAT_DRHO_PREV2 = 23
POL_STAT_DU_PREV2 = 24
POL_STAT_FLUID = 25
POL_STAT_FLUID_RHO_NAN = 26
POL_STAT_FLUID_RHO_INF = 27
POL_STAT_FLUID_RHO_MIN = 28
POL_STAT_FLUID_RHO_MAX = 29
POL_STAT_FLUID_U_NAN = 30
POL_STAT_FLUID_U_INF = 31
POL_STAT_FLUID_UMAX = 32
POL_STAT_FLUID_POP_NAN = 33
POL_STAT_FLUID_POP_INF = 34
POL_STAT_INT_CELLS = 35
POL_STAT_INT_WATER = 36
POL_STAT_INT_AIR = 37
POL_STAT_INT_UMAX = 38
POL_STAT_INT_UMIN = 39
POL_STAT_INT_USUM = 40
POL_STAT_SUM_UMAG_ACTIVE = 41
POL_STAT_UMAG_ACTIVE = 42
POL_NSTAT = 43
STAT_SLOTS = ("cells", "sum_drho", "sum_drho2", "max_abs_drho", "mom_x", "mom_y", "mom_z", "sum_umag", "sum_umag2", "max_umag", "sum_mach_active", "active", "mach_gt", "sum_ux", "sum_uy", "sum_uz", "sum_ux2", "sum_uy2", "sum_uz2", "sum_umag_top", "sum_umag_mid", "sum_umag_bot", "outlet_down", "drho_prev2", "du_prev2", "fluid", "fluid_rho_nan", "fluid_rho_inf", "fluid_rho_min", "fluid_rho_max", "fluid_u_nan", "fluid_u_inf", "fluid_umax", "fluid_pop_nan", "fluid_pop_inf", "int_cells", "int_water", "int_air", "int_umax", "int_umin", "int_usum", "sum_umag_active", "umag_active")

POL_THERMAL_CONVECTION = 0x1
POL_THERMAL_RESET_SOURCE = 0x2
POL_MP_APPLY = 0x1
POL_PRESSURE_SURFACE_TENSION = 0x4

STORAGE_CODES = {"f32": POL_F32, "f16": POL_F16, "bf16": POL_BF16, "f64": POL_F64}


class PolDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("storage", C.c_int32), ("options", C.c_uint32), ("device", C.c_int32),
        ("tau", C.c_double), ("cs_smag", C.c_double), ("nu_sgs_max", C.c_double),
        ("gravity", C.c_double * 3), ("darcy_K", C.c_double), ("forch_CF", C.c_double),
        ("inlet_rho", C.c_double), ("inlet_u", C.c_double * 3), ("outlet_rho", C.c_double),
    ]


class PolBuffers(C.Structure):
    _fields_ = [
        ("f", C.c_void_p * 2), ("cell_flags", C.c_void_p), ("force", C.c_void_p),
        ("perm", C.c_void_p), ("rho", C.c_void_p), ("u", C.c_void_p),
        ("active_chunks", C.c_void_p), ("n_chunks", C.c_uint32), ("chunk_cells", C.c_uint32),
        ("particle_drag", C.c_void_p),
        ("chunk_records", C.c_void_p), ("stored_cells", C.c_uint32), ("layout_flags", C.c_uint32),
        ("nu_sgs", C.c_void_p), ("les_phase", C.c_void_p),
    ]


class PolParticleDesc(C.Structure):
    _fields_ = [
        ("count", C.c_int32), ("drag_law", C.c_int32),
        ("rho_fluid", C.c_double), ("nu_fluid", C.c_double), ("drag_max_per_mass", C.c_double),
        ("relax_alpha", C.c_double), ("gravity", C.c_double * 3), ("dt", C.c_double),
        ("max_disp", C.c_double), ("wall_damping", C.c_double),
        ("reaction", C.c_int32), ("reserved", C.c_int32),
    ]


class PolParticleBuffers(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "pos", "vel", "radius", "mass", "active", "drag_new", "drag_old", "drag", "u_fluid", "re_p", "cd",
        "drag_field", "u_field", "reaction_field")]


class PolStatsOpts(C.Structure):
    _fields_ = [("prev_rho", C.c_void_p), ("prev_u", C.c_void_p), ("phase", C.c_void_p), ("z_top", C.c_int32),
                ("z_mid", C.c_int32), ("z_outlet", C.c_int32), ("with_populations", C.c_int32)]


class PolThermalDesc(C.Structure):
    _fields_ = [("omega", C.c_double), ("options", C.c_uint32), ("reserved", C.c_int32)]


class PolThermalBuffers(C.Structure):
    _fields_ = [("g", C.c_void_p * 2), ("temperature", C.c_void_p * 2), ("heat_source", C.c_void_p),
                ("velocity", C.c_void_p), ("conductivity", C.c_void_p), ("heat_source_base", C.c_void_p)]


class PolMultiphaseDesc(C.Structure):
    _fields_ = [("sigma", C.c_double), ("mobility", C.c_double), ("interface_width", C.c_double),
                ("rho_water", C.c_double), ("rho_air", C.c_double), ("dt", C.c_double)]


class PolMultiphaseBuffers(C.Structure):
    _fields_ = [("phi", C.c_void_p * 2), ("phi_new", C.c_void_p), ("mu", C.c_void_p), ("phase", C.c_void_p)]


class PolInterfaceFields(C.Structure):
    _fields_ = [("grad_phi", C.c_void_p), ("grad_mu", C.c_void_p), ("normal", C.c_void_p), ("curvature", C.c_void_p),
                ("surface_force", C.c_void_p)]


class PolPourDesc(C.Structure):
    _fields_ = [("center_x", C.c_double), ("center_y", C.c_double), ("radius", C.c_double),
                ("stream_height", C.c_double), ("velocity", C.c_double), ("flow_rate", C.c_double),
                ("dt", C.c_double), ("height", C.c_int32), ("reserved", C.c_int32)]


class PolAdjointBuffers(C.Structure):
    _fields_ = [("lam", C.c_void_p * 2), ("grad_perm", C.c_void_p), ("rev_flags", C.c_void_p)]


# name -> (restype, argtypes); every symbol declared in the public header
_P = C.c_void_p
SIGNATURES = {
    "pol_version": (C.c_int, []),
    "pol_status_string": (C.c_char_p, [C.c_int]),
    "pol_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "pol_create": (C.c_int, [C.POINTER(_P), C.POINTER(PolDesc)]),
    "pol_destroy": (C.c_int, [_P]),
    "pol_last_error": (C.c_char_p, [_P]),
    "pol_bind": (C.c_int, [_P, C.POINTER(PolBuffers)]),
    "pol_set_params": (C.c_int, [_P, C.POINTER(PolDesc)]),
    "pol_current": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "pol_set_current": (C.c_int, [_P, C.c_int]),
    "pol_init_equilibrium": (C.c_int, [_P, _P, _P, _P]),
    "pol_step": (C.c_int, [_P, C.c_int, _P]),
    "pol_step_ex": (C.c_int, [_P, C.c_int, C.c_uint32, _P, _P]),
    "pol_step_from": (C.c_int, [_P, _P, _P, C.c_uint32, _P, _P]),
    "pol_stream_only": (C.c_int, [_P, C.c_int, _P]),
    "pol_macroscopic": (C.c_int, [_P, C.c_uint32, _P, _P]),
    "pol_reduce": (C.c_int, [_P, _P, _P]),
    "pol_reduce_noalloc": (C.c_int, [C.POINTER(PolDesc), C.POINTER(PolBuffers), _P, _P]),
    "pol_les_viscosity": (C.c_int, [_P, _P, _P]),
    "pol_particles_bind": (C.c_int, [_P, C.POINTER(PolParticleDesc), C.POINTER(PolParticleBuffers)]),
    "pol_particles_couple": (C.c_int, [_P, _P]),
    "pol_particles_compute_coupling": (C.c_int, [_P, _P, _P]),
    "pol_particles_relax": (C.c_int, [_P, C.c_double, _P]),
    "pol_particles_clear_reaction": (C.c_int, [_P, _P]),
    "pol_add_reaction_forces": (C.c_int, [_P, _P, _P]),
    "pol_particles_advance": (C.c_int, [_P, _P]),
    "pol_coupled_step": (C.c_int, [_P, C.c_int, C.c_uint32, _P, _P]),
    "pol_particle_index": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P, _P, _P]),
    "pol_adjoint_bind": (C.c_int, [_P, C.POINTER(PolAdjointBuffers)]),
    "pol_adjoint_step": (C.c_int, [_P, _P, C.c_int, _P]),
    "pol_adjoint_reset": (C.c_int, [_P, _P]),
    "pol_adjoint_sweep": (C.c_int, [_P, _P, C.c_uint64, C.c_int, C.c_int, _P]),
    "pol_step_chain": (C.c_int, [_P, _P, C.c_uint64, C.c_int, _P]),
    "pol_adjoint_collect": (C.c_int, [_P, _P, _P]),
    "pol_field_stats": (C.c_int, [_P, C.POINTER(PolStatsOpts), _P, _P, _P]),
    "pol_export_fields": (C.c_int, [_P, _P, _P, _P]),
    "pol_stabilize": (C.c_int, [_P, C.c_double, C.c_double, C.c_double, _P]),
    "pol_thermal_bind": (C.c_int, [_P, C.POINTER(PolThermalDesc), C.POINTER(PolThermalBuffers)]),
    "pol_thermal_set": (C.c_int, [_P, C.POINTER(PolThermalDesc)]),
    "pol_thermal_init": (C.c_int, [_P, _P]),
    "pol_thermal_step": (C.c_int, [_P, C.c_int, _P, _P]),
    "pol_thermal_heat_flux": (C.c_int, [_P, _P, _P]),
    "pol_thermal_stats": (C.c_int, [_P, _P, _P]),
    "pol_thermal_boundary": (C.c_int, [_P, C.c_int, _P, C.c_double, C.c_double, _P]),
    "pol_multiphase_bind": (C.c_int, [_P, C.POINTER(PolMultiphaseDesc), C.POINTER(PolMultiphaseBuffers)]),
    "pol_multiphase_chemical_potential": (C.c_int, [_P, _P, _P]),
    "pol_multiphase_interface": (C.c_int, [_P, C.c_uint32, C.POINTER(PolInterfaceFields), _P]),
    "pol_multiphase_step": (C.c_int, [_P, _P]),
    "pol_multiphase_density": (C.c_int, [_P, _P]),
    "pol_multiphase_current": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "pol_pour_force": (C.c_int, [_P, C.POINTER(PolPourDesc), _P]),
    "pol_pour_phase": (C.c_int, [_P, C.POINTER(PolPourDesc), _P, _P]),
    "pol_density_drive": (C.c_int, [_P, _P, C.c_double, C.c_int, C.c_double, C.c_double, C.c_double, _P]),
    "pol_pressure_force": (C.c_int, [_P, _P, C.c_int, C.c_double, C.c_double, _P]),
    "pol_filter_blockage": (C.c_int, [_P, _P, C.c_int32, _P, _P, _P]),
    "pol_filter_block_particles": (C.c_int, [_P, _P, _P, _P, _P, _P, C.c_int32, C.c_double, C.c_uint32, _P]),
}

_lib = None


def load() -> C.CDLL:
    """Load and type the library (once).  Raises ImportError if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("POUROVER_HIP_LIB") or LIB_PATH)
    if not path.exists():
        raise ImportError(
            "libpourover_hip.so not found at %s: build it with __graft_entry__.build() "
            "(there is no CPU fallback for the HIP hot path)" % path)
    lib = C.CDLL(str(path))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.pol_version() != POL_ABI_VERSION:
        raise ImportError("libpourover_hip.so ABI %d != %d" % (lib.pol_version(), POL_ABI_VERSION))
    _lib = lib
    return lib


class PolError(RuntimeError):
    """Non-zero status from the C ABI."""

    def __init__(self, status: int, message: str):
        self.status = status
        super().__init__(message)


def check(status: int, ctx=None, what: str = "") -> None:
    if status == POL_OK:
        return
    lib = load()
    detail = ""
    if ctx:
        detail = (lib.pol_last_error(ctx) or b"").decode(errors="replace")
    base = lib.pol_status_string(status).decode()
    raise PolError(status, "%s failed: %s%s" % (what or "pol call", base, (" - " + detail) if detail else ""))


def stream_handle(stream=None):
    """hipStream_t of a torch stream (default: current stream)."""
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)
