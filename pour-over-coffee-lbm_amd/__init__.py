"""MI355X-native D3Q19 lattice-Boltzmann engine (pour-over coffee brewer).

Hot path: hand-written gfx950 HIP kernels in csrc/ (fused stream-collide,
particle coupling, adjoint, field diagnostics, thermal D3Q7, phase field,
body-force producers), driven through the C ABI of include/pourover_hip.h
via ctypes.  Host layer mirrors the reference's backend plugin API
(backends.py) and solver protocol (solver.py).

Importing the package needs numpy/torch only; the HIP library is loaded on
first use and raises ImportError if it has not been built (no CPU fallback).
"""
from . import lattice  # noqa: F401

__all__ = ["lattice", "geometry", "engine", "solver", "backends", "fields", "workloads", "particles", "adjoint",
           "diagnostics", "thermal", "multiphase", "forcing"]
__version__ = "0.1.0"


def load_library():
    """Load libpourover_hip.so (ImportError if missing)."""
    from . import _lib

    return _lib.load()
