"""ctypes front-end of the C oracle (oracle/lbm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the reference of comparison and by bench.py's cpu_baseline leg -- never by the product.

Arrays follow the device layout: populations [19, nz, ny, nx], scalar fields
[nz, ny, nx], vector fields [3, nz, ny, nx], particle vectors [3, P].
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle_lbm.so"

# descriptor/flag constants of include/pourover_hip.h (plain copies: the
# oracle must not import the product package)
POL_PERIODIC_X, POL_PERIODIC_Y, POL_PERIODIC_Z = 0x1, 0x2, 0x4
POL_USE_FLAGS, POL_LES, POL_FORCE_FIELD, POL_PERM_FIELD, POL_GRAVITY = 0x8, 0x10, 0x20, 0x40, 0x80
POL_PARTICLE_FIELD = 0x100
POL_LES_FD = 0x200
NRED = 12


class Desc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("storage", C.c_int32), ("options", C.c_uint32), ("device", C.c_int32),
        ("tau", C.c_double), ("cs_smag", C.c_double), ("nu_sgs_max", C.c_double),
        ("gravity", C.c_double * 3), ("darcy_K", C.c_double), ("forch_CF", C.c_double),
        ("inlet_rho", C.c_double), ("inlet_u", C.c_double * 3), ("outlet_rho", C.c_double),
    ]


class ParticleParams(C.Structure):
    pass


def _pp_fields(real):
    return [("count", C.c_int), ("drag_law", C.c_int), ("rho_fluid", real), ("nu_fluid", real),
            ("drag_max_per_mass", real),
            ("alpha", real), ("g", real * 3), ("dt", real), ("max_disp", real), ("wall_damping", real)]


class ParticleParams64(C.Structure):
    _fields_ = _pp_fields(C.c_double)


class ParticleParams32(C.Structure):
    _fields_ = _pp_fields(C.c_float)


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            from . import build_oracle  # type: ignore

            build_oracle.build()
        _lib = C.CDLL(str(LIB))
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def make_desc(nx, ny, nz, *, tau, options=0, cs_smag=0.0, nu_sgs_max=0.1, gravity=(0.0, 0.0, 0.0), darcy_K=1e-3,
              forch_CF=0.0, inlet_rho=1.0, inlet_u=(0.0, 0.0, 0.0), outlet_rho=1.0) -> Desc:
    d = Desc()
    d.nx, d.ny, d.nz = nx, ny, nz
    d.storage = 3
    d.options = options
    d.device = 0
    d.tau = tau
    d.cs_smag = cs_smag
    d.nu_sgs_max = nu_sgs_max
    d.gravity = (C.c_double * 3)(*gravity)
    d.darcy_K = darcy_K
    d.forch_CF = forch_CF
    d.inlet_rho = inlet_rho
    d.inlet_u = (C.c_double * 3)(*inlet_u)
    d.outlet_rho = outlet_rho
    return d


def desc_from_dict(cfg: dict) -> Desc:
    """Build a Desc from the keyword dict the product's workloads use."""
    keys = ("tau", "options", "cs_smag", "nu_sgs_max", "gravity", "darcy_K", "forch_CF", "inlet_rho", "inlet_u",
            "outlet_rho")
    return make_desc(cfg["nx"], cfg["ny"], cfg["nz"], **{k: cfg[k] for k in keys if k in cfg})


def _dtype(prec):
    return np.float64 if prec == "f64" else np.float32


def _sfx(prec):
    return "_f64" if prec == "f64" else "_f32"


def step(desc: Desc, f: np.ndarray, flags=None, force=None, perm=None, want_macro=False, prec="f64", pdrag=None,
         nu=None):
    """One oracle step.  Returns (f_new, rho, u, red) (rho/u None unless want_macro).
    pdrag: [4, nz, ny, nx] particle-drag field (used when desc has POL_PARTICLE_FIELD).
    nu: [nz, ny, nx] given eddy viscosity (used when desc has POL_LES and POL_LES_FD)."""
    dt = _dtype(prec)
    f = np.ascontiguousarray(f, dtype=dt)
    fout = np.zeros_like(f)
    shape = f.shape[1:]
    rho = np.empty(shape, dtype=dt) if want_macro else None
    u = np.empty((3,) + shape, dtype=dt) if want_macro else None
    red = np.zeros(NRED, dtype=np.float64)
    fl = np.ascontiguousarray(flags, dtype=np.uint32) if flags is not None else None
    fo = np.ascontiguousarray(force, dtype=dt) if force is not None else None
    pm = np.ascontiguousarray(perm, dtype=dt) if perm is not None else None
    pd = np.ascontiguousarray(pdrag, dtype=dt) if pdrag is not None else None
    nv = np.ascontiguousarray(nu, dtype=dt) if nu is not None else None
    fn = getattr(lib(), "oracle_step_nu" + _sfx(prec))
    fn.restype = C.c_int
    rc = fn(C.byref(desc), _p(fl), _p(fo), _p(pm), _p(pd), _p(nv), _p(f), _p(fout), _p(rho), _p(u), _p(red))
    if rc != 0:
        raise RuntimeError("oracle_step failed")
    return fout, rho, u, red


def run(desc: Desc, f: np.ndarray, nsteps: int, flags=None, force=None, perm=None, prec="f64"):
    """nsteps oracle steps; returns (f, rho, u, red) of the last step."""
    rho = u = red = None
    for i in range(nsteps):
        f, rho, u, red = step(desc, f, flags, force, perm, want_macro=(i == nsteps - 1), prec=prec)
    return f, rho, u, red


def moments(desc: Desc, f: np.ndarray, flags=None, force=None, perm=None, prec="f64", pdrag=None):
    """Moments (rho, u, red) of the state f (the gather of the next step)."""
    _, rho, u, red = step(desc, f, flags, force, perm, want_macro=True, prec=prec, pdrag=pdrag)
    return rho, u, red


def init_equilibrium(nx, ny, nz, rho=None, u=None, prec="f64") -> np.ndarray:
    dt = _dtype(prec)
    f = np.empty((19, nz, ny, nx), dtype=dt)
    r = np.ascontiguousarray(rho, dtype=dt) if rho is not None else None
    uu = np.ascontiguousarray(u, dtype=dt) if u is not None else None
    getattr(lib(), "oracle_init_equilibrium" + _sfx(prec))(C.c_int(nx), C.c_int(ny), C.c_int(nz), _p(r), _p(uu), _p(f))
    return f


def cpu_backend_step(f, f_new, rho, u, solid, tau, prec="f32"):
    """CPUBackend.execute_collision_streaming restatement (in place)."""
    nz, ny, nx = f.shape[1:]
    real = C.c_double if prec == "f64" else C.c_float
    getattr(lib(), "oracle_cpu_backend_step" + _sfx(prec))(
        C.c_int(nx), C.c_int(ny), C.c_int(nz), real(tau), _p(f), _p(f_new), _p(rho), _p(u), _p(solid))


def particle_index(pos: np.ndarray, dims):
    """coffee_particles.py:1159-1174 index/fraction in float32; pos [3, n]."""
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    n = pos.shape[1]
    idx = np.empty((3, n), dtype=np.int32)
    frac = np.empty((3, n), dtype=np.float32)
    nx, ny, nz = dims
    lib().oracle_particle_index_f64(_p(pos), C.c_int(n), C.c_int(nx), C.c_int(ny), C.c_int(nz), _p(idx), _p(frac))
    return idx, frac


def drag_coefficient(re: float, prec="f64") -> float:
    fn = getattr(lib(), "oracle_drag_coefficient" + _sfx(prec))
    real = C.c_double if prec == "f64" else C.c_float
    fn.restype = real
    fn.argtypes = [real]
    return float(fn(re))


def particle_params(count, prec="f64", **kw):
    cls = ParticleParams64 if prec == "f64" else ParticleParams32
    pp = cls()
    pp.count = count
    for k, v in kw.items():
        if k == "g":
            pp.g = (type(pp.g)._type_ * 3)(*v)
        else:
            setattr(pp, k, v)
    return pp


def particles_couple(dims, pp, st: dict, u_field, force, prec="f64"):
    """st: dict of pos, vel, radius, mass, active, drag_old (+ outputs created)."""
    dt = _dtype(prec)
    P = pp.count
    for k in ("u_fluid", "drag_new", "drag"):
        st.setdefault(k, np.zeros((3, P), dtype=dt))
    for k in ("re_p", "cd"):
        st.setdefault(k, np.zeros(P, dtype=dt))
    nx, ny, nz = dims
    getattr(lib(), "oracle_particles_couple" + _sfx(prec))(
        C.c_int(nx), C.c_int(ny), C.c_int(nz), C.byref(pp), _p(st["pos"]), _p(st["vel"]), _p(st["radius"]),
        _p(st["mass"]), _p(st["active"]), _p(u_field), _p(force), _p(st["u_fluid"]), _p(st["drag_new"]),
        _p(st["drag_old"]), _p(st["drag"]), _p(st["re_p"]), _p(st["cd"]))


def particles_compute(dims, pp, st: dict, u_field, field, explicit=True, prec="f64"):
    """compute_two_way_coupling_forces without relaxation: field is the [3][cells]
    reaction field (explicit, accumulated: clear it first) or the [4][cells]
    drag field (implicit)."""
    dt = _dtype(prec)
    P = pp.count
    for k in ("u_fluid", "drag_new"):
        st.setdefault(k, np.zeros((3, P), dtype=dt))
    for k in ("re_p", "cd"):
        st.setdefault(k, np.zeros(P, dtype=dt))
    nx, ny, nz = dims
    getattr(lib(), "oracle_particles_compute" + _sfx(prec))(
        C.c_int(nx), C.c_int(ny), C.c_int(nz), C.byref(pp), _p(st["pos"]), _p(st["vel"]), _p(st["radius"]),
        _p(st["mass"]), _p(st["active"]), _p(u_field), _p(field), _p(st["u_fluid"]), _p(st["drag_new"]),
        _p(st["re_p"]), _p(st["cd"]), C.c_int(1 if explicit else 0))


def particles_relax(st: dict, alpha: float, prec="f64"):
    """apply_under_relaxation on st's drag_new / drag_old (creates drag)."""
    dt = _dtype(prec)
    P = st["drag_new"].shape[1]
    st.setdefault("drag", np.zeros((3, P), dtype=dt))
    real = C.c_double if prec == "f64" else C.c_float
    getattr(lib(), "oracle_particles_relax" + _sfx(prec))(
        C.c_int(P), real(alpha), _p(st["active"]), _p(st["drag_new"]), _p(st["drag_old"]), _p(st["drag"]))


def add_reaction_forces(force: np.ndarray, reaction: np.ndarray, flags=None) -> None:
    """legacy/lbm_solver.py:1478-1483: force += reaction on non-solid cells (in place)."""
    if flags is None:
        force += reaction
        return
    kind = (np.asarray(flags, dtype=np.uint32) >> 24) & 0x7
    fluid = kind != 1
    force += np.where(fluid[None], reaction, 0.0)


def particles_advance(dims, pp, st: dict, flags, force, prec="f64"):
    nx, ny, nz = dims
    fl = np.ascontiguousarray(flags, dtype=np.uint32) if flags is not None else None
    getattr(lib(), "oracle_particles_advance" + _sfx(prec))(
        C.c_int(nx), C.c_int(ny), C.c_int(nz), C.byref(pp), _p(fl), _p(st["pos"]), _p(st["vel"]), _p(st["mass"]),
        _p(st["active"]), _p(st["u_fluid"]), _p(st["drag_new"]), _p(force))
