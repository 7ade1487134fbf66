"""Host-side driver of one HIP solver context (device buffers + C ABI calls).

Torch-ROCm tensors are only the device allocations; every computation on
them is a launch of libpourover_hip.so through the C ABI (_lib.py).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib as L
from .layout import build_compact_layout

POL_CHUNK = L.POL_CHUNK
WRITE_GRANULE = 128  # bytes: HBM writes narrower than this line cost a read-modify-write
# Population layout, chunk size and run alignment by storage, from the
# 224^3 V60 step (ms per step):
#   dense layout, 16 / 32 / 64-cell chunks: FP32 0.174 / 0.144 / 0.159,
#     16-bit 0.114 / 0.098 / 0.110, FP64 0.255 (32)
#   compact, runs packed (align 1): FP32 0.166 / 0.154 / 0.141,
#     16-bit 0.091 / 0.093 / 0.100, FP64 0.294 (16)
#   compact, (chunk, align) = (8, 8) / (16, 16) / (32, 32), first measured:
#     FP32 0.125 / 0.124 / 0.127, FP16 0.081 / 0.082 / 0.087,
#     BF16 - / 0.090 / -, FP64 0.239 / 0.249 / -
#   and again with the current kernels (two runs each; the collisionless
#   pass in brackets; (16, 8) and (32, 16) slower for every storage):
#     FP32 0.1151, 0.1128 (0.1118) / 0.1169, 0.1150 (0.1147) / 0.1204, 0.1185
#     FP16 0.0642, 0.0654 (0.0610) / 0.0670, 0.0673 (0.0621) / 0.0749, 0.0738
#     BF16 0.0666, 0.0653 (0.0618) / 0.0689, 0.0674 (0.0626) / 0.0771, 0.0744
#   and with the launched pieces read as first cells plus one record per
#   row (4 / 2 / 8 cells, chunk = align, two runs each, same box):
#     FP32 0.1096, 0.1075 (0.1065) / 0.1217, 0.1197 / 0.1120, 0.1102 (0.1090)
#     FP16 0.0619, 0.0622, 0.0639, 0.0637 / 0.0681, 0.0665 /
#          0.0627, 0.0637, 0.0624, 0.0651
#     BF16 0.0641, 0.0639, 0.0642, 0.0648 / 0.0689, 0.0677 /
#          0.0666, 0.0649, 0.0647, 0.0658
# Aligned runs write whole aligned segments and leave 1-6 % padding where
# the dense chunks leave 23 % solid lanes (stored slots per plane at 224^3:
# 3 783 296 for 4-cell runs, 3 845 120 for 8, 3 962 688 for 16).
COMPACT_DEFAULT = {"f32": True, "f16": True, "bf16": True, "f64": True}
CHUNK_CELLS = {"f32": 32, "f16": 32, "bf16": 32, "f64": 16}
CHUNK_CELLS_COMPACT = {"f32": 4, "f16": 4, "bf16": 4, "f64": 8}
COMPACT_ALIGN = {"f32": 4, "f16": 4, "bf16": 4, "f64": 8}
_STORAGE_TORCH = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16, "f64": torch.float64}


@dataclass
class SolverConfig:
    """Descriptor of one lattice (fields of pol_desc plus buffer options)."""

    nx: int
    ny: int
    nz: int
    precision: str = "f32"           # population storage: f32 | f16 | bf16 | f64
    tau: float = 0.6
    periodic: Sequence[bool] = (True, True, True)
    les: bool = False
    cs_smag: float = 0.17
    nu_sgs_max: float = 0.1          # les_turbulence.py:375
    gravity: Optional[Sequence[float]] = None
    darcy_K: float = 1e-3
    forch_CF: float = 0.55
    inlet_rho: float = 1.0
    inlet_u: Sequence[float] = (0.0, 0.0, 0.0)
    outlet_rho: float = 1.0
    force_field: bool = False
    perm_field: bool = False
    particle_field: bool = False     # implicit particle-drag field (two-way coupling)
    les_fd: bool = False             # with les: the reference's finite-difference eddy viscosity
    device: int = 0

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def real_dtype(self) -> torch.dtype:
        return torch.float64 if self.precision == "f64" else torch.float32

    @property
    def storage_dtype(self) -> torch.dtype:
        return _STORAGE_TORCH[self.precision]

    @property
    def bytes_per_population(self) -> int:
        return {"f32": 4, "f16": 2, "bf16": 2, "f64": 8}[self.precision]

    def options(self, use_flags: bool) -> int:
        o = 0
        px, py, pz = self.periodic
        o |= L.POL_PERIODIC_X if px else 0
        o |= L.POL_PERIODIC_Y if py else 0
        o |= L.POL_PERIODIC_Z if pz else 0
        o |= L.POL_USE_FLAGS if use_flags else 0
        o |= L.POL_LES if self.les else 0
        o |= L.POL_FORCE_FIELD if self.force_field else 0
        o |= L.POL_PERM_FIELD if self.perm_field else 0
        o |= L.POL_GRAVITY if self.gravity is not None else 0
        o |= L.POL_PARTICLE_FIELD if self.particle_field else 0
        o |= L.POL_LES_FD if self.les_fd else 0
        return o

    def desc(self, use_flags: bool) -> L.PolDesc:
        d = L.PolDesc()
        d.nx, d.ny, d.nz = self.nx, self.ny, self.nz
        d.storage = L.STORAGE_CODES[self.precision]
        d.options = self.options(use_flags)
        d.device = self.device
        d.tau = self.tau
        d.cs_smag = self.cs_smag
        d.nu_sgs_max = self.nu_sgs_max
        d.gravity = (C.c_double * 3)(*(self.gravity or (0.0, 0.0, 0.0)))
        d.darcy_K = self.darcy_K
        d.forch_CF = self.forch_CF
        d.inlet_rho = self.inlet_rho
        d.inlet_u = (C.c_double * 3)(*self.inlet_u)
        d.outlet_rho = self.outlet_rho
        return d

    def oracle_dict(self, use_flags: bool) -> dict:
        """Keyword form of the same descriptor (for tests)."""
        return dict(nx=self.nx, ny=self.ny, nz=self.nz, tau=self.tau, options=self.options(use_flags),
                    cs_smag=self.cs_smag, nu_sgs_max=self.nu_sgs_max,
                    gravity=tuple(self.gravity or (0.0, 0.0, 0.0)), darcy_K=self.darcy_K, forch_CF=self.forch_CF,
                    inlet_rho=self.inlet_rho, inlet_u=tuple(self.inlet_u), outlet_rho=self.outlet_rho)


def _as_int32_view(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32))


class HipContext:
    """One pol_ctx with its device buffers."""

    def __init__(self, cfg: SolverConfig, flags: Optional[np.ndarray] = None, chunk_cells: Optional[int] = None,
                 compact: Optional[bool] = None, align: Optional[int] = None):
        """compact: store populations in the row-run layout (layout.py), on
        by default whenever cell flags are given; chunk_cells / align override
        the per-storage launch chunk and run alignment."""
        if not torch.cuda.is_available():
            raise RuntimeError("HipContext needs a visible HIP device (torch.cuda.is_available() is False)")
        self.lib = L.load()
        self.config = cfg
        self.device = torch.device("cuda", cfg.device)
        shape = (cfg.nz, cfg.ny, cfg.nx)
        self.shape = shape
        self.cells = cfg.cells
        sd, rd = cfg.storage_dtype, cfg.real_dtype
        self.rho = torch.ones(shape, dtype=rd, device=self.device)
        self.u = torch.zeros((3,) + shape, dtype=rd, device=self.device)
        self.flags = _as_int32_view(flags).to(self.device) if flags is not None else None
        self.compact = (flags is not None and COMPACT_DEFAULT[cfg.precision]) if compact is None else bool(compact)
        table = CHUNK_CELLS_COMPACT if self.compact else CHUNK_CELLS
        self.chunk_cells = int(chunk_cells or table[cfg.precision])
        self.align = int(align or COMPACT_ALIGN[cfg.precision]) if self.compact else 1
        if self.compact and flags is None:
            raise ValueError("the compact layout needs cell flags")
        self._set_layout(flags)
        # zero = rest state f = w
        self.f = [torch.zeros((19, self.plane), dtype=sd, device=self.device) for _ in range(2)]
        self.force = torch.zeros((3,) + shape, dtype=rd, device=self.device) if cfg.force_field else None
        self.perm = torch.full(shape, cfg.darcy_K, dtype=rd, device=self.device) if cfg.perm_field else None
        self.pdrag = torch.zeros((4,) + shape, dtype=rd, device=self.device) if cfg.particle_field else None
        # finite-difference LES: the eddy viscosity and the optional phase field
        # of its interface suppression (None: phi = 1)
        self.nu_sgs = torch.zeros(shape, dtype=rd, device=self.device) if cfg.les_fd else None
        self.les_phase: Optional[torch.Tensor] = None
        self._ctx = C.c_void_p()
        self._use_flags = flags is not None
        st = self.lib.pol_create(C.byref(self._ctx), C.byref(cfg.desc(self._use_flags)))
        L.check(st, None, "pol_create")
        self._bind()

    # -- lifecycle -----------------------------------------------------------
    def _bind(self) -> None:
        L.check(self.lib.pol_bind(self._ctx, C.byref(self.buffers())), self._ctx, "pol_bind")

    def _bind_keep_state(self) -> None:
        """Re-bind the buffers (after one of the optional ones changed) keeping
        the current-state index."""
        cur = self.current
        self._bind()
        self.set_current(cur)

    def buffers(self) -> L.PolBuffers:
        """The pol_buffers of this context (f[0] / f[1] are the two ping-pong
        buffers; the current state is f[self.current])."""
        b = L.PolBuffers()
        b.f[0] = self.f[0].data_ptr()
        b.f[1] = self.f[1].data_ptr()
        b.cell_flags = self.flags.data_ptr() if self.flags is not None else None
        b.force = self.force.data_ptr() if self.force is not None else None
        b.perm = self.perm.data_ptr() if self.perm is not None else None
        b.rho = self.rho.data_ptr()
        b.u = self.u.data_ptr()
        if self.records is not None:
            b.chunk_records = self.records.data_ptr()
            b.active_chunks = self.chunks.data_ptr()  # the launched pieces
            b.n_chunks = self.layout.n_chunks
            b.stored_cells = self.layout.stored_cells
            b.layout_flags = 0 if self.layout.x_face else L.POL_LAYOUT_NO_X_FACE
            b.chunk_cells = self.chunk_cells
        elif self.chunks is not None:
            b.active_chunks = self.chunks.data_ptr()
            b.n_chunks = self.chunks.numel()
            b.chunk_cells = self.chunk_cells
        b.particle_drag = self.pdrag.data_ptr() if self.pdrag is not None else None
        b.nu_sgs = self.nu_sgs.data_ptr() if self.nu_sgs is not None else None
        b.les_phase = self.les_phase.data_ptr() if self.les_phase is not None else None
        return b

    def _set_layout(self, flags: Optional[np.ndarray]) -> None:
        """Launch list and population layout for these flags: the compact
        row-run layout (its pieces and row records), or the dense layout with
        (when flags are given) a dense chunk list."""
        cfg = self.config
        if self.compact:
            self.layout = build_compact_layout(flags, tuple(bool(p) for p in cfg.periodic), self.chunk_cells,
                                               self.align)
            self.records = _as_int32_view(self.layout.row_records.reshape(-1)).to(self.device)
            self.chunks = _as_int32_view(self.layout.piece_words).to(self.device)
            self.plane = self.layout.stored_cells
        else:
            self.layout = None
            self.records = None
            self.chunks = self._chunk_list(flags)
            self.plane = cfg.cells

    def dense_populations(self, t: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Populations [19, nz, ny, nx] of a state buffer (default: the current
        one) in the dense layout; cells outside the stored runs read 0 (= w)."""
        t = self.f_current if t is None else t
        if self.layout is None:
            return t.view((19,) + self.shape)
        idx = torch.from_numpy(self.layout.dense_index).to(self.device)
        slots = torch.from_numpy(self.layout.run_slots).to(self.device)
        out = torch.zeros((19, self.cells), dtype=t.dtype, device=self.device)
        out[:, idx] = t[:, slots]
        return out.view((19,) + self.shape)

    def load_dense_populations(self, dense: torch.Tensor, t: Optional[torch.Tensor] = None) -> None:
        """Write dense populations [19, nz, ny, nx] into a state buffer."""
        t = self.f_current if t is None else t
        d = dense.to(self.device, t.dtype).reshape(19, -1)
        if self.layout is None:
            t.copy_(d)
            return
        idx = torch.from_numpy(self.layout.dense_index).to(self.device)
        slots = torch.from_numpy(self.layout.run_slots).to(self.device)
        t.zero_()
        t[:, slots] = d[:, idx]

    def _chunk_list(self, flags: Optional[np.ndarray]) -> Optional[torch.Tensor]:
        """Launch list: first cells of the chunk_cells-cell row chunks holding
        any non-solid cell (None when every chunk does or there are no flags).
        Chunks of solid cells get no threads; solid lanes of a listed chunk
        store zero deviations so written granules are whole.  When a chunk is
        narrower than the write granule, the chunks sharing a granule are listed
        together (rows of a multiple of the granule's cells)."""
        if flags is None:
            return None
        cfg = self.config
        ch = self.chunk_cells
        kind = (np.asarray(flags, dtype=np.uint32) >> L.POL_KIND_SHIFT) & L.POL_KIND_MASK
        open_cell = kind != L.POL_KIND_SOLID
        nch = -(-cfg.nx // ch)
        pad = nch * ch - cfg.nx
        if pad:
            open_cell = np.concatenate([open_cell, np.zeros(open_cell.shape[:2] + (pad,), bool)], axis=2)
        live = open_cell.reshape(cfg.nz, cfg.ny, nch, ch).any(axis=3)
        per_granule = WRITE_GRANULE // (cfg.bytes_per_population * ch)
        if not pad and per_granule > 1 and (cfg.nx // ch) % per_granule == 0:
            grp = live.reshape(-1, per_granule).any(axis=1, keepdims=True)
            live = np.broadcast_to(grp, (grp.shape[0], per_granule)).reshape(live.shape)
        zz, yy, cc = np.nonzero(live)
        if zz.size == 0 or zz.size == live.size:
            return None
        base = ((zz.astype(np.int64) * cfg.ny + yy) * cfg.nx + cc * ch).astype(np.uint32)
        return _as_int32_view(base).to(self.device)

    def close(self) -> None:
        if self._ctx:
            self.lib.pol_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._ctx

    # -- state -----------------------------------------------------------------
    @property
    def current(self) -> int:
        i = C.c_int(0)
        L.check(self.lib.pol_current(self._ctx, C.byref(i)), self._ctx, "pol_current")
        return i.value

    def set_current(self, i: int) -> None:
        L.check(self.lib.pol_set_current(self._ctx, int(i)), self._ctx, "pol_set_current")

    @property
    def f_current(self) -> torch.Tensor:
        return self.f[self.current]

    @property
    def f_next(self) -> torch.Tensor:
        return self.f[1 - self.current]

    def set_flags(self, flags: np.ndarray) -> None:
        if not self._use_flags:
            raise ValueError("context was created without cell flags")
        cur = self.current
        dense = [self.dense_populations(t) for t in self.f] if self.compact else None
        self.flags.copy_(_as_int32_view(flags).to(self.device))
        self._set_layout(flags)
        if dense is not None:
            # the stored runs follow the flags: remap both states
            sd = self.config.storage_dtype
            self.f = [torch.zeros((19, self.plane), dtype=sd, device=self.device) for _ in range(2)]
            for t, d in zip(self.f, dense):
                self.load_dense_populations(d, t)
        self._bind()
        self.set_current(cur)

    def synchronize(self) -> None:
        torch.cuda.synchronize(self.device)

    def set_params(self, **changes) -> None:
        """Change physics scalars (tau, cs_smag, gravity, darcy_K, forch_CF, inlet_u, ...)."""
        from dataclasses import replace

        cfg = replace(self.config, **changes)
        if (cfg.options(self._use_flags) & ~L.POL_LES) != (self.config.options(self._use_flags) & ~L.POL_LES):
            raise ValueError("set_params cannot switch options other than LES (gravity/fields/periodicity)")
        L.check(self.lib.pol_set_params(self._ctx, C.byref(cfg.desc(self._use_flags))), self._ctx, "pol_set_params")
        self.config = cfg

    # -- launches --------------------------------------------------------------
    def init_equilibrium(self, rho: Optional[torch.Tensor] = None, u: Optional[torch.Tensor] = None) -> None:
        rd = self.config.real_dtype
        r = rho.to(self.device, rd).contiguous() if rho is not None else None
        uu = u.to(self.device, rd).contiguous() if u is not None else None
        st = self.lib.pol_init_equilibrium(self._ctx, L.ptr(r), L.ptr(uu), L.stream_handle())
        L.check(st, self._ctx, "pol_init_equilibrium")
        if r is not None or uu is not None:
            torch.cuda.current_stream().synchronize()  # r/uu temporaries die here

    def step(self, n: int = 1, what: int = 0) -> Optional[np.ndarray]:
        red = np.zeros(L.POL_NRED, dtype=np.float64) if what & L.POL_WANT_REDUCE else None
        st = self.lib.pol_step_ex(self._ctx, int(n), int(what),
                                  red.ctypes.data_as(C.c_void_p) if red is not None else None, L.stream_handle())
        L.check(st, self._ctx, "pol_step")
        return red

    def stream_only(self, n: int = 1) -> None:
        """n collisionless streaming passes (pol_stream_only): the step's gather
        and stores without its collision, the layout's bandwidth ceiling."""
        L.check(self.lib.pol_stream_only(self._ctx, int(n), L.stream_handle()), self._ctx, "pol_stream_only")

    def step_from(self, fin: torch.Tensor, fout: torch.Tensor, what: int = 0) -> Optional[np.ndarray]:
        red = np.zeros(L.POL_NRED, dtype=np.float64) if what & L.POL_WANT_REDUCE else None
        st = self.lib.pol_step_from(self._ctx, L.ptr(fin), L.ptr(fout), int(what),
                                    red.ctypes.data_as(C.c_void_p) if red is not None else None, L.stream_handle())
        L.check(st, self._ctx, "pol_step_from")
        return red

    def macroscopic(self, macro: bool = True, reduce: bool = False) -> Optional[np.ndarray]:
        what = (L.POL_WANT_MACRO if macro else 0) | (L.POL_WANT_REDUCE if reduce else 0)
        red = np.zeros(L.POL_NRED, dtype=np.float64) if reduce else None
        st = self.lib.pol_macroscopic(self._ctx, what, red.ctypes.data_as(C.c_void_p) if red is not None else None,
                                      L.stream_handle())
        L.check(st, self._ctx, "pol_macroscopic")
        return red

    def reduce(self) -> dict:
        red = self.macroscopic(macro=False, reduce=True)
        return reduction_dict(red)


def reduction_dict(red: np.ndarray) -> dict:
    return {k: float(v) for k, v in zip(L.RED_SLOTS, red)}
