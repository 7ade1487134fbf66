"""Field wrappers with the Taichi-field surface callers of the reference use.

Callers of the reference read solver fields with ``to_numpy()``,
``from_numpy()``, ``fill()`` and ``shape`` (main.py:1074-1076,
src/visualization/lbm_diagnostics.py:234-235).  The HIP engine keeps fields
as device tensors in [..., z, y, x] order (x fastest); these wrappers present
the reference's index order [i, j, k] = [x, y, z] on the host side:
  scalar field     -> numpy (NX, NY, NZ)
  vector field     -> numpy (NX, NY, NZ, 3)   (ti.Vector.field(3))
  population field -> numpy (Q, NX, NY, NZ)   (legacy 4-D f[q, i, j, k])
"""
from __future__ import annotations

from typing import Callable, Optional

import numpy as np
import torch

from .lattice import W

# Device population buffers hold D_q = SCALE (f_q - w_q) (exact integer
# weights 36 w_q in the kernels' arithmetic).
SCALE = 36.0


def decode_populations(t: torch.Tensor) -> np.ndarray:
    """Raw populations f [19, ...] (host, float64 for FP64 storage else float32)."""
    if t.dtype == torch.float64:
        return t.cpu().numpy() / SCALE + W.reshape((-1,) + (1,) * (t.dim() - 1))
    a = t.float().cpu().numpy()
    return (a / np.float32(SCALE) + W.astype(np.float32).reshape((-1,) + (1,) * (t.dim() - 1))).astype(np.float32)


def encode_populations(f: np.ndarray) -> np.ndarray:
    """Inverse of decode_populations (float64 host array)."""
    f = np.asarray(f, dtype=np.float64)
    return SCALE * (f - W.reshape((-1,) + (1,) * (f.ndim - 1)))


class DeviceField:
    """View of a device tensor shaped [C..., nz, ny, nx] as a reference field."""

    def __init__(self, tensor: torch.Tensor, kind: str = "scalar", before_read: Optional[Callable[[], None]] = None,
                 after_write: Optional[Callable[[], None]] = None, deviation: bool = False, ctx=None):
        assert kind in ("scalar", "vector", "population")
        self._t = tensor
        self.kind = kind
        self._before_read = before_read
        self._after_write = after_write
        self._deviation = deviation  # population storage holds f - w
        # populations: the HipContext that maps its storage layout (dense or
        # compact row runs) to [19, nz, ny, nx]
        self._ctx = ctx

    # -- shape / dtype -----------------------------------------------------
    @property
    def shape(self):
        if self._ctx is not None:
            nz, ny, nx = self._ctx.shape
            return (self._t.shape[0], nx, ny, nz)
        nz, ny, nx = self._t.shape[-3:]
        if self.kind == "scalar":
            return (nx, ny, nz)
        if self.kind == "vector":
            return (nx, ny, nz)
        return (self._t.shape[0], nx, ny, nz)

    @property
    def n(self):
        return 3 if self.kind == "vector" else None

    @property
    def dtype(self):
        return self._t.dtype

    def to_torch(self) -> torch.Tensor:
        """The underlying device tensor in device layout ([C,] nz, ny, nx; a
        population buffer in its storage layout, [19, slots])."""
        if self._before_read:
            self._before_read()
        return self._t

    # -- host transfer --------------------------------------------------------
    def to_numpy(self) -> np.ndarray:
        if self._before_read:
            self._before_read()
        t = self._t if self._ctx is None else self._ctx.dense_populations(self._t)
        if self._deviation:
            a = decode_populations(t)
        else:
            a = t.cpu().numpy()
        if self.kind == "scalar":
            return np.ascontiguousarray(a.transpose(2, 1, 0))
        if self.kind == "vector":
            return np.ascontiguousarray(a.transpose(3, 2, 1, 0))
        return np.ascontiguousarray(a.transpose(0, 3, 2, 1))

    def from_numpy(self, arr) -> None:
        a = np.asarray(arr)
        if self.kind == "scalar":
            dev = a.transpose(2, 1, 0)
        elif self.kind == "vector":
            dev = a.transpose(3, 2, 1, 0)
        else:
            dev = a.transpose(0, 3, 2, 1)
            if self._deviation:
                dev = encode_populations(dev)
        src = torch.from_numpy(np.ascontiguousarray(dev)).to(self._t.dtype)
        if self._ctx is not None:
            self._ctx.load_dense_populations(src, self._t)
        else:
            self._t.copy_(src.to(self._t.device))
        if self._after_write:
            self._after_write()

    def fill(self, value) -> None:
        if self.kind == "population" and self._deviation:
            w = torch.tensor(W, dtype=torch.float64, device=self._t.device).view((-1,) + (1,) * (self._t.dim() - 1))
            self._t.copy_((SCALE * (float(value) - w)).to(self._t.dtype).expand_as(self._t))
        elif self.kind == "vector" and not np.isscalar(value):
            v = torch.tensor(list(value), dtype=self._t.dtype, device=self._t.device).view(3, 1, 1, 1)
            self._t.copy_(v.expand_as(self._t))
        else:
            self._t.fill_(value)
        if self._after_write:
            self._after_write()

    def copy_from(self, other: "DeviceField") -> None:
        self._t.copy_(other.to_torch())
        if self._after_write:
            self._after_write()

    def __getitem__(self, idx):
        """Element access in reference order (host round trip; debugging only)."""
        return self.to_numpy()[idx]

    def __repr__(self):
        return "DeviceField(kind=%s, shape=%s, dtype=%s)" % (self.kind, self.shape, self._t.dtype)
