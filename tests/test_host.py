"""Host logic without a GPU: backend registry and error codes, protocol
coverage, population encoding, seeded workloads, the launch chunk list and
the one-GPU guard of bench.py."""
from __future__ import annotations

import inspect
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest
import torch

from pourover_amd import _lib as L
from pourover_amd import backends as B
from pourover_amd import fields as F
from pourover_amd import geometry as G
from pourover_amd import workloads as WL
from pourover_amd.adjoint import auto_checkpoint
from pourover_amd.engine import CHUNK_CELLS, WRITE_GRANULE, HipContext
from pourover_amd.solver import HipLBMSolver, LBMSolverProtocol, ParticleCoupledLBMProtocol

ROOT = Path(__file__).resolve().parents[1]


# ---- backend registry (compute_backends.py:35-65,70-75,758-783) -------------------
@pytest.mark.parametrize("name", ["apple", "cuda", "cpu"])
def test_reference_only_backends_raise_import_failed(name):
    with pytest.raises(B.BackendInitializationError) as e:
        B.ComputeBackendFactory().create_backend(name, use_cache=False)
    assert e.value.error_code == "IMPORT_FAILED"


def test_unknown_backend_code():
    with pytest.raises(B.BackendInitializationError) as e:
        B.ComputeBackendFactory().create_backend("metal-x", use_cache=False)
    assert e.value.error_code == "UNKNOWN_BACKEND"


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-device path")
def test_hip_backend_without_device_is_platform_unavailable():
    with pytest.raises(B.BackendInitializationError) as e:
        B.ComputeBackendFactory().create_backend("hip", use_cache=False)
    assert e.value.error_code == "PLATFORM_UNAVAILABLE"
    with pytest.raises(RuntimeError):
        HipContext(WL.taylor_green_config(8))


def test_status_codes_map_to_reference_errors():
    for status, cls, code in ((L.POL_ERR_ALLOC, B.MemoryAllocationError, "ALLOCATION_FAILED"),
                              (L.POL_ERR_NO_DEVICE, B.BackendInitializationError, "PLATFORM_UNAVAILABLE"),
                              (L.POL_ERR_INVALID, B.ComputeExecutionError, "EXECUTION_FAILED"),
                              (L.POL_ERR_HIP, B.ComputeExecutionError, "EXECUTION_FAILED")):
        with pytest.raises(cls) as e:
            B.raise_for_status(L.PolError(status, "x"), "hip", "op")
        assert e.value.error_code == code
        assert isinstance(e.value, B.BackendError)


def test_platform_type_has_hip():
    assert B.PlatformType.AMD_HIP.value == "hip"


# ---- solver protocol (lbm_protocol.py:28-257) -------------------------------------
def test_solver_implements_every_protocol_member():
    members = [n for n in dir(ParticleCoupledLBMProtocol) if not n.startswith("_")]
    members += list(getattr(LBMSolverProtocol, "__annotations__", {}))
    for name in set(members):
        assert hasattr(HipLBMSolver, name) or name in inspect.signature(HipLBMSolver.__init__).parameters \
            or name in ("f", "f_new", "rho", "u", "solid", "phase"), name
    for extra in ("init_fields", "clear_body_force", "add_particle_reaction_forces", "step_with_two_way_coupling",
                  "get_velocity_field_for_thermal_coupling", "swap_fields"):
        assert callable(getattr(HipLBMSolver, extra)), extra


# ---- population encoding (stored D = 36 (f - w)) ----------------------------------
def test_population_encoding_round_trip():
    rng = np.random.default_rng(1)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    f = w[:, None] * (1 + 0.1 * rng.standard_normal((19, 64)))
    d = F.encode_populations(f)
    assert np.allclose(d, 36 * (f - w[:, None]))
    back = F.decode_populations(torch.from_numpy(d))
    assert np.allclose(back, f, rtol=0, atol=1e-15)
    # rest state encodes to exactly zero
    assert np.array_equal(F.encode_populations(np.repeat(w[:, None], 3, axis=1)), np.zeros((19, 3)))


# ---- descriptor bits -------------------------------------------------------------
def test_config_options():
    wl = WL.v60_config(32, "f16", particle_field=True, perm_field=True)
    o = wl.config.options(True)
    for bit in (L.POL_USE_FLAGS, L.POL_LES, L.POL_GRAVITY, L.POL_PERM_FIELD, L.POL_PARTICLE_FIELD):
        assert o & bit
    assert not o & (L.POL_PERIODIC_X | L.POL_PERIODIC_Y | L.POL_PERIODIC_Z | L.POL_FORCE_FIELD)
    d = wl.config.desc(True)
    assert d.storage == L.POL_F16 and d.nx == 32 and abs(d.gravity[2] + 1e-5) < 1e-20
    tg = WL.taylor_green_config(16)
    assert tg.options(False) == L.POL_PERIODIC_X | L.POL_PERIODIC_Y | L.POL_PERIODIC_Z


# ---- seeded inputs ----------------------------------------------------------------
def test_seed_particles_deterministic_and_inside_fluid():
    geo = G.build_v60(G.V60Params.scaled(48))
    a = WL.seed_particles(geo, 2000)
    b = WL.seed_particles(geo, 2000)
    assert np.array_equal(a.pos, b.pos) and np.array_equal(a.radius, b.radius)
    ci = a.pos.astype(np.int64)
    assert (geo.kind[ci[2], ci[1], ci[0]] == L.POL_KIND_FLUID).all()
    assert not geo.porous[ci[2], ci[1], ci[0]].any()
    d = 2 * a.radius
    assert abs(np.median(d) / 1.5 - 1) < 0.05
    assert abs(np.std(np.log(d)) / 0.4 - 1) < 0.1
    assert np.allclose(a.mass, 1.2 * np.pi * d**3 / 6)


def test_permeability_field_range():
    geo = G.build_v60(G.V60Params.scaled(48))
    K = WL.permeability_field(geo)
    k = K[geo.porous]
    assert k.min() >= 1e-4 and k.max() <= 1e-2
    assert np.array_equal(K, WL.permeability_field(geo))
    # log-uniform: mean of log10 K ~ -3
    assert abs(np.log10(k).mean() + 3) < 0.1


# ---- launch chunk list ------------------------------------------------------------
class _Cfg:
    def __init__(self, cfg, chunk_cells=None):
        self.config = cfg
        self.device = "cpu"
        self.chunk_cells = chunk_cells or CHUNK_CELLS[cfg.precision]


def _chunks(cfg, flags, chunk_cells=None):
    c = HipContext._chunk_list(_Cfg(cfg, chunk_cells), flags)
    return None if c is None else c.numpy().view(np.uint32)


@pytest.mark.parametrize("prec,ch", [("f32", None), ("f16", None), ("f64", None), ("f16", 16), ("f32", 8),
                                     ("f32", 32)])
def test_chunk_list_covers_every_open_cell(prec, ch):
    wl = WL.v60_config(64, prec)
    flags = wl.geometry.flags
    ch = ch or CHUNK_CELLS[prec]
    base = _chunks(wl.config, flags, ch)
    assert base is not None and (np.diff(base.astype(np.int64)) > 0).all()
    assert (base % ch == 0).all()
    covered = np.zeros(flags.size, bool)
    for c in base:
        covered[c:c + ch] = True
    open_cell = (wl.geometry.kind != L.POL_KIND_SOLID).ravel()
    assert covered[open_cell].all()
    bpp = wl.config.bytes_per_population
    per = WRITE_GRANULE // (bpp * ch)
    if per > 1:
        # the chunks sharing a 64-byte granule are listed together
        _, cnt = np.unique(base // (ch * per), return_counts=True)
        assert (cnt == per).all()
    else:
        # only chunks that hold fluid, each a whole number of granules
        assert (ch * bpp) % WRITE_GRANULE == 0
        assert all(open_cell[c:c + ch].any() for c in base)


def test_chunk_list_absent_for_open_box():
    flags, _ = G.periodic_box((32, 32, 32))
    assert _chunks(WL.taylor_green_config(32), flags) is None


# ---- bench contract ---------------------------------------------------------------
def test_bench_refuses_more_than_one_gpu():
    p = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert p.returncode != 0
    assert "one GPU" in p.stderr


# ---- adjoint state store ------------------------------------------------------------
def test_auto_checkpoint_keeps_states_that_fit():
    mb = 1 << 20
    # 500 states of 56 MB fit in half of 200 GB: keep them all
    assert auto_checkpoint(500, 56 * mb, 200 * 1024 * mb) is None
    # otherwise the largest divisor of steps not above 50
    assert auto_checkpoint(500, 56 * mb, 20 * 1024 * mb) == 50
    assert auto_checkpoint(120, 56 * mb, 1024 * mb) == 40
    assert auto_checkpoint(97, 56 * mb, 1024 * mb) == 1
    assert auto_checkpoint(30, 56 * mb, 1024 * mb) == 30
