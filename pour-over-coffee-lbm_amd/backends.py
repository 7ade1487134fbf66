"""Compute-backend plugin layer with the new ``'hip'`` backend.

Mirrors the reference's backend API so callers written against it find the
same names, signatures and error behaviour:
  error hierarchy           src/core/backends/compute_backends.py:35-65
  PlatformType              compute_backends.py:70-75 (+ AMD_HIP = "hip")
  ComputeBackend ABC        compute_backends.py:338-539
  concrete-backend contract cpu_backend.py:14-25,192-237,298-352
                            (execute_collision_streaming(memory_adapter, params),
                             get_backend_info, estimate_memory_usage, validate_platform)
  ComputeBackendFactory     compute_backends.py:544-905 (singleton, lock, cache,
                            string dispatch, fallback chain)
The reference's Taichi backends ('apple', 'cuda', 'cpu') are not part of this
build; asking the factory for them raises BackendInitializationError
('IMPORT_FAILED'), an unknown name raises it with 'UNKNOWN_BACKEND', exactly
where the reference raises (compute_backends.py:758-783).
"""
from __future__ import annotations

import logging
import threading
import time
from abc import ABC, abstractmethod
from contextlib import contextmanager
from enum import Enum
from typing import Any, Callable, Dict, List, Optional

import numpy as np

from . import _lib as L


# ---- errors ----------------------------------------------------------------
class BackendError(Exception):
    def __init__(self, message: str, backend_type: Optional[str] = None, error_code: Optional[str] = None):
        self.backend_type = backend_type or "unknown"
        self.error_code = error_code or "UNKNOWN_ERROR"
        super().__init__(message)


class PlatformDetectionError(BackendError):
    pass


class BackendInitializationError(BackendError):
    pass


class ComputeExecutionError(BackendError):
    pass


class MemoryAllocationError(BackendError):
    pass


class PerformanceDegradationError(BackendError):
    pass


def raise_for_status(err: "L.PolError", backend_type: str = "hip", op: str = "") -> None:
    """Map a C-ABI status onto the reference's error classes."""
    st = err.status
    msg = "%s: %s" % (op, err) if op else str(err)
    if st == L.POL_ERR_ALLOC:
        raise MemoryAllocationError(msg, backend_type, "ALLOCATION_FAILED") from err
    if st == L.POL_ERR_NO_DEVICE:
        raise BackendInitializationError(msg, backend_type, "PLATFORM_UNAVAILABLE") from err
    raise ComputeExecutionError(msg, backend_type, "EXECUTION_FAILED") from err


class PlatformType(Enum):
    APPLE_SILICON = "apple"
    NVIDIA_CUDA = "cuda"
    CPU_REFERENCE = "cpu"
    AMD_HIP = "hip"
    UNKNOWN = "unknown"


class BackendPriority(Enum):
    CRITICAL = 1
    HIGH = 2
    MEDIUM = 3
    LOW = 4


class BackendLogger:
    def __init__(self):
        self.logger = logging.getLogger("compute_backends")

    def info(self, message: str, backend_type: Optional[str] = None):
        self.logger.info("%s%s", "[%s] " % backend_type if backend_type else "", message)

    def warning(self, message: str, backend_type: Optional[str] = None):
        self.logger.warning("%s%s", "[%s] " % backend_type if backend_type else "", message)

    def error(self, message: str, backend_type: Optional[str] = None):
        self.logger.error("%s%s", "[%s] " % backend_type if backend_type else "", message)


backend_logger = BackendLogger()


# ---- platform detection -----------------------------------------------------
class PlatformDetector:
    """HIP detection: torch built for ROCm, a visible device, and the C ABI."""

    def __init__(self):
        self._lock = threading.Lock()
        self.detected_platforms: List[Dict[str, Any]] = []

    def hip_available(self) -> bool:
        try:
            import torch

            if not getattr(torch.version, "hip", None):
                return False
            if not torch.cuda.is_available():
                return False
            n = L.C.c_int(0)
            return L.load().pol_device_count(L.C.byref(n)) == L.POL_OK and n.value > 0
        except Exception:
            return False

    def detect_all_platforms(self) -> List[Dict[str, Any]]:
        with self._lock:
            out = [{"platform_type": PlatformType.AMD_HIP, "backend_name": "hip", "display_name": "AMD HIP (gfx950)",
                    "priority": BackendPriority.HIGH, "available": self.hip_available()}]
            self.detected_platforms = out
            return out

    def get_optimal_platform(self) -> Dict[str, Any]:
        for p in self.detect_all_platforms():
            if p["available"]:
                return p
        return {"platform_type": PlatformType.UNKNOWN, "backend_name": "hip", "display_name": "none",
                "priority": BackendPriority.LOW, "available": False}

    def validate_platform_availability(self, backend_type: str) -> bool:
        if backend_type == "hip":
            return self.hip_available()
        return False


platform_detector = PlatformDetector()


# ---- backend ABC --------------------------------------------------------------
class ComputeBackend(ABC):
    """Backend base class (compute_backends.py:338-539)."""

    def __init__(self, backend_type: str = "base"):
        self.backend_type = backend_type
        self.performance_metrics: Dict[str, Any] = {}
        self.error_history: List[Dict[str, Any]] = []
        self.is_initialized = False
        self.creation_time = time.time()
        self.last_error = None
        self.execution_times: List[float] = []
        self.memory_usage: List[float] = []
        self.operation_count = 0

    @abstractmethod
    def execute_collision_streaming(self, memory_adapter, params: Optional[Dict[str, Any]] = None, **kwargs) -> None:
        ...

    @abstractmethod
    def apply_boundary_conditions(self, memory_adapter, params: Optional[Dict[str, Any]] = None, **kwargs) -> None:
        ...

    @abstractmethod
    def compute_macroscopic_quantities(self, memory_adapter, params: Optional[Dict[str, Any]] = None,
                                       **kwargs) -> None:
        ...

    @abstractmethod
    def get_platform_info(self) -> Dict[str, Any]:
        ...

    @abstractmethod
    def get_performance_metrics(self) -> Dict[str, Any]:
        ...

    def initialize_backend(self) -> None:
        try:
            self._perform_initialization()
            self.is_initialized = True
        except BackendError:
            raise
        except Exception as e:
            msg = "%s backend initialization failed: %s" % (self.backend_type, e)
            self._record_error(msg, "INIT_FAILED")
            raise BackendInitializationError(msg, self.backend_type, "INIT_FAILED") from e

    def cleanup_backend(self) -> None:
        try:
            self._perform_cleanup()
        finally:
            self.is_initialized = False

    @contextmanager
    def safe_execution(self, operation_name: str):
        start = time.time()
        try:
            yield
        except ComputeExecutionError:
            raise
        except Exception as e:
            msg = "%s failed: %s" % (operation_name, e)
            self._record_error(msg, "EXECUTION_FAILED")
            raise ComputeExecutionError(msg, self.backend_type, "EXECUTION_FAILED") from e
        finally:
            self.execution_times.append(time.time() - start)
            self.operation_count += 1
            if len(self.execution_times) > 1000:
                self.execution_times = self.execution_times[-500:]

    def _perform_initialization(self) -> None:
        pass

    def _perform_cleanup(self) -> None:
        pass

    def _record_error(self, error_msg: str, error_code: str) -> None:
        rec = {"timestamp": time.time(), "message": error_msg, "code": error_code, "backend_type": self.backend_type}
        self.error_history.append(rec)
        self.last_error = rec
        if len(self.error_history) > 100:
            self.error_history = self.error_history[-50:]

    def get_error_summary(self) -> Dict[str, Any]:
        return {"total_errors": len(self.error_history), "last_error": self.last_error,
                "error_rate": len(self.error_history) / max(1, self.operation_count),
                "backend_type": self.backend_type}

    def get_basic_metrics(self) -> Dict[str, Any]:
        if not self.execution_times:
            return {"status": "no_data"}
        return {"backend_type": self.backend_type, "total_operations": self.operation_count,
                "avg_execution_time": float(np.mean(self.execution_times)),
                "min_execution_time": float(np.min(self.execution_times)),
                "max_execution_time": float(np.max(self.execution_times)),
                "uptime_seconds": time.time() - self.creation_time, "is_initialized": self.is_initialized,
                "error_summary": self.get_error_summary()}


def _params_of(params, kwargs) -> Dict[str, Any]:
    """Accept both call forms of the reference: (adapter, params_dict) and (adapter, **kwargs)."""
    out: Dict[str, Any] = {}
    if params:
        out.update(params)
    out.update(kwargs)
    return out


class HipBackend(ComputeBackend):
    """MI355X backend: every operation is a launch of libpourover_hip.so.

    The memory adapter is a HipMemoryAdapter (solver.py), which owns the HIP
    context and the device buffers; the attribute contract f / f_new / rho /
    u / solid of cpu_backend.py:207-212 is kept on it.
    """

    def __init__(self):
        super().__init__("hip")
        self.performance_metrics = {"collision_time": 0.0, "streaming_time": 0.0, "boundary_time": 0.0,
                                    "total_time": 0.0, "cells": 0, "steps": 0}

    def _perform_initialization(self) -> None:
        if not platform_detector.hip_available():
            raise BackendInitializationError("no HIP device visible", "hip", "PLATFORM_UNAVAILABLE")

    def validate_platform(self) -> bool:
        return platform_detector.hip_available()

    @staticmethod
    def _ctx(memory_adapter):
        ctx = getattr(memory_adapter, "hip_context", None)
        if ctx is None:
            raise ComputeExecutionError("memory adapter carries no HIP context (use HipMemoryAdapter)", "hip",
                                        "EXECUTION_FAILED")
        return ctx

    def execute_collision_streaming(self, memory_adapter, params: Optional[Dict[str, Any]] = None, **kwargs) -> None:
        """One (or params['steps']) fused stream-collide step(s).

        Collision, streaming and wall treatment are one launch, so the three
        timers of cpu_backend.py:218-237 collapse into collision_time.
        """
        p = _params_of(params, kwargs)
        steps = int(p.get("steps", 1))
        ctx = self._ctx(memory_adapter)
        if "tau" in p and abs(float(p["tau"]) - ctx.config.tau) > 1e-12:
            raise ComputeExecutionError("tau %.6g differs from the context's %.6g; rebuild the solver to change it"
                                        % (float(p["tau"]), ctx.config.tau), "hip", "EXECUTION_FAILED")
        t0 = time.perf_counter()
        with self.safe_execution("execute_collision_streaming"):
            try:
                ctx.step(steps)
                if p.get("sync", True):
                    ctx.synchronize()
            except L.PolError as e:
                raise_for_status(e, "hip", "execute_collision_streaming")
        dt = time.perf_counter() - t0
        m = self.performance_metrics
        m["collision_time"] = dt
        m["streaming_time"] = 0.0
        m["boundary_time"] = 0.0
        m["total_time"] = dt
        m["cells"] = ctx.cells
        m["steps"] = steps

    def apply_boundary_conditions(self, memory_adapter, params: Optional[Dict[str, Any]] = None, **kwargs) -> None:
        """Walls, inlet and outlet are applied inside the fused step; nothing to launch."""
        self._ctx(memory_adapter)
        self.performance_metrics["boundary_time"] = 0.0

    def compute_macroscopic_quantities(self, memory_adapter, params: Optional[Dict[str, Any]] = None,
                                       **kwargs) -> None:
        ctx = self._ctx(memory_adapter)
        with self.safe_execution("compute_macroscopic_quantities"):
            try:
                ctx.macroscopic()
            except L.PolError as e:
                raise_for_status(e, "hip", "compute_macroscopic_quantities")

    def get_platform_info(self) -> Dict[str, Any]:
        import torch

        info: Dict[str, Any] = {"platform": "AMD HIP", "arch": "gfx950", "torch": torch.__version__,
                                "hip_runtime": getattr(torch.version, "hip", None)}
        if torch.cuda.is_available():
            prop = torch.cuda.get_device_properties(0)
            info.update({"device_name": prop.name, "total_memory_gb": prop.total_memory / 2**30,
                         "multiprocessors": prop.multi_processor_count})
        return info

    def _calculate_throughput(self) -> float:
        m = self.performance_metrics
        if m["total_time"] > 0 and m["cells"]:
            return m["cells"] * m["steps"] / 1e6 / m["total_time"]
        return 0.0

    def get_performance_metrics(self) -> Dict[str, Any]:
        m = dict(self.performance_metrics)
        m["throughput_mlups"] = self._calculate_throughput()
        return m

    def get_backend_info(self) -> Dict[str, Any]:
        return {"name": "HIP Backend", "type": "AMD MI355X (gfx950)", "gpu_acceleration": True,
                "precision": "f32 / f16 / bf16 / f64 population storage", "memory_layout": "SoA f[q][z][y][x]",
                "optimization_level": "fused stream-collide", "debugging_support": False}

    def estimate_memory_usage(self, nx: int, ny: int, nz: int, bytes_per_pop: int = 4) -> float:
        """GB for two population buffers + flags + rho/u (cf. cpu_backend.py:342-348)."""
        cells = nx * ny * nz
        return (2 * 19 * bytes_per_pop * cells + 4 * cells + 16 * cells) / 1024**3


# ---- factory -----------------------------------------------------------------
_REGISTRY: Dict[str, Callable[[], ComputeBackend]] = {"hip": HipBackend}
_REFERENCE_ONLY = ("apple", "cuda", "cpu")


def register_backend(name: str, ctor: Callable[[], ComputeBackend]) -> None:
    _REGISTRY[name] = ctor


class ComputeBackendFactory:
    """Thread-safe singleton factory (compute_backends.py:544-905)."""

    _instance = None
    _lock = threading.Lock()

    def __new__(cls):
        if cls._instance is None:
            with cls._lock:
                if cls._instance is None:
                    cls._instance = super().__new__(cls)
                    cls._instance._initialized = False
        return cls._instance

    def __init__(self):
        if getattr(self, "_initialized", False):
            return
        self.backend_cache: Dict[str, ComputeBackend] = {}
        self.creation_history: List[Dict[str, Any]] = []
        self.fallback_chain = ["hip"]
        self.lock = threading.Lock()
        self._initialized = True

    def create_optimal_backend(self, prefer_performance: bool = True) -> ComputeBackend:
        with self.lock:
            best = platform_detector.get_optimal_platform()
            name = best["backend_name"]
            try:
                backend = self._create_backend_with_validation(name)
                self._record(name)
                return backend
            except Exception:
                return self._attempt_fallback_creation(exclude=[name])

    def create_backend(self, backend_type: str, validate: bool = True, use_cache: bool = True) -> ComputeBackend:
        with self.lock:
            if use_cache and backend_type in self.backend_cache:
                cached = self.backend_cache[backend_type]
                if cached.is_initialized:
                    return cached
                del self.backend_cache[backend_type]
            try:
                backend = self._create_backend_with_validation(backend_type, validate)
            except BackendInitializationError as e:
                # keep the specific code (UNKNOWN_BACKEND, IMPORT_FAILED, ...)
                raise BackendInitializationError("creating %s backend failed: %s" % (backend_type, e),
                                                 backend_type, e.error_code) from e
            except Exception as e:
                raise BackendInitializationError("creating %s backend failed: %s" % (backend_type, e),
                                                 backend_type, "CREATION_FAILED") from e
            if use_cache:
                self.backend_cache[backend_type] = backend
            self._record(backend_type)
            return backend

    def _create_backend_with_validation(self, backend_type: str, validate: bool = True) -> ComputeBackend:
        if backend_type in _REFERENCE_ONLY:
            raise BackendInitializationError("%s backend is a Taichi backend of the reference and is not part of "
                                             "this package" % backend_type, backend_type, "IMPORT_FAILED")
        ctor = _REGISTRY.get(backend_type)
        if ctor is None:
            raise BackendInitializationError("unknown backend type: %s" % backend_type, backend_type,
                                             "UNKNOWN_BACKEND")
        if validate and not platform_detector.validate_platform_availability(backend_type):
            raise BackendInitializationError("%s platform unavailable" % backend_type, backend_type,
                                             "PLATFORM_UNAVAILABLE")
        backend = ctor()
        if validate:
            backend.initialize_backend()
        return backend

    def _attempt_fallback_creation(self, exclude: Optional[List[str]] = None) -> ComputeBackend:
        exclude = exclude or []
        for name in self.fallback_chain:
            if name in exclude:
                continue
            try:
                return self._create_backend_with_validation(name)
            except Exception as e:  # try the next one
                backend_logger.warning("fallback to %s failed: %s" % (name, e))
        raise BackendInitializationError("no compute backend available", "all", "ALL_BACKENDS_FAILED")

    def get_available_backends(self) -> List[Dict[str, Any]]:
        out = []
        for p in platform_detector.detect_all_platforms():
            out.append({**p, "is_available": bool(p["available"]),
                        "validation_status": "passed" if p["available"] else "failed"})
        return out

    def _record(self, backend_type: str) -> None:
        self.creation_history.append({"timestamp": time.time(), "backend_type": backend_type, "status": "success"})
        if len(self.creation_history) > 100:
            self.creation_history = self.creation_history[-50:]

    def get_factory_statistics(self) -> Dict[str, Any]:
        return {"total_backends_created": len(self.creation_history), "cached_backends": list(self.backend_cache),
                "creation_history": self.creation_history[-10:], "fallback_chain": list(self.fallback_chain)}

    def cleanup_all_backends(self) -> None:
        with self.lock:
            for b in self.backend_cache.values():
                b.cleanup_backend()
            self.backend_cache.clear()


backend_factory = ComputeBackendFactory()


def create_optimal_backend(**kwargs) -> ComputeBackend:
    return backend_factory.create_optimal_backend(**kwargs)


def create_backend(backend_type: str, **kwargs) -> ComputeBackend:
    return backend_factory.create_backend(backend_type, **kwargs)


def get_available_backends() -> List[Dict[str, Any]]:
    return backend_factory.get_available_backends()


def cleanup_backends() -> None:
    backend_factory.cleanup_all_backends()
