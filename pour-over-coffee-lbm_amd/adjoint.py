"""Differentiable V60 run: outlet flux gradient by a hand-written adjoint.

Objective (lattice units): J = sum over the last `window` steps of the
outlet mass flux, sum over outlet cells of -rho u_z (the moments each step
collides; the reduction slot POL_RED_OUTFLUX).  Parameters: the per-cell
permeability K of the filter layer and the inlet velocity u_in.

The reverse sweep runs csrc/adjoint.hip one step at a time, newest first.
Forward states come from one of two stores:
  - checkpoint=None: every forward state is kept (the 128^3 cup in the
    compact layout: 56 MB per FP32 state, 28 GB for 500 steps), the forward
    run writes them as it goes and the reverse sweep recomputes nothing: a
    reverse step costs one adjoint kernel;
  - checkpoint=K: a state every K steps; a segment is re-run forward from
    its checkpoint into a segment store (one fused step per state, the
    segment in one native call, pol_step_chain) and then swept in reverse
    (pol_adjoint_sweep), so memory holds the checkpoints plus one segment,
    and a reverse step costs one forward step more.  (Rebuilding the next
    segment on a second stream beside the sweep was measured slower: the
    adjoint kernel fills the device, 0.0637 against 0.0602 ms per reverse
    step at 128^3.)
The default, checkpoint="auto", keeps every state when they fit in half of
the free device memory (on 288 GB of HBM that covers the 128^3 cup for
thousands of steps) and checkpoints every 50 steps (or the largest divisor
of `steps` below that) otherwise.

There is no adjoint in the reference (no autograd or checkpointing exists in
src/); OutletFluxFunction wraps the sweep as a torch.autograd.Function so
callers can treat the whole run as a differentiable op.
"""
from __future__ import annotations

import ctypes as C
import time
from typing import Optional, Sequence, Union

import numpy as np
import torch

from . import _lib as L
from .engine import HipContext
from .workloads import V60Workload


def auto_checkpoint(steps: int, state_bytes: int, free_bytes: int) -> Optional[int]:
    """None (keep every state) when steps states fit in half of free_bytes,
    else the largest divisor of steps not above 50."""
    if steps * state_bytes <= free_bytes // 2:
        return None
    return max(d for d in range(1, min(50, steps) + 1) if steps % d == 0)


class AdjointProblem:
    def __init__(self, wl: V60Workload, K: np.ndarray, steps: int = 500, window: int = 100,
                 checkpoint: Union[int, None, str] = "auto", compact: Optional[bool] = None):
        cfg = wl.config
        if not cfg.perm_field:
            raise ValueError("the adjoint workload needs SolverConfig(perm_field=True)")
        if cfg.precision not in ("f32", "f64"):
            raise ValueError("the adjoint runs with f32 or f64 populations")
        self.wl = wl
        self.ctx = HipContext(cfg, wl.geometry.flags, compact=compact)
        if checkpoint == "auto":
            state_bytes = self.ctx.f[0].numel() * self.ctx.f[0].element_size()
            free, _ = torch.cuda.mem_get_info(self.ctx.device)
            checkpoint = auto_checkpoint(steps, state_bytes, free)
        self.store_all = checkpoint is None
        if self.store_all:
            checkpoint = steps
        if steps % checkpoint:
            raise ValueError("steps must be a multiple of the checkpoint interval")
        self.steps, self.window, self.checkpoint = steps, window, checkpoint
        self.ctx.perm.copy_(torch.as_tensor(K, dtype=cfg.real_dtype))
        shape = self.ctx.shape
        rd = cfg.real_dtype
        dev = self.ctx.device
        # adjoint populations in the forward state's storage layout
        self.lam = [torch.zeros(tuple(self.ctx.f[0].shape), dtype=rd, device=dev) for _ in range(2)]
        self.grad_perm = torch.zeros(shape, dtype=torch.float64, device=dev)
        self.rev_flags = torch.from_numpy(np.ascontiguousarray(wl.geometry.rev_flags).view(np.int32)).to(dev)
        b = L.PolAdjointBuffers()
        b.lam[0] = self.lam[0].data_ptr()
        b.lam[1] = self.lam[1].data_ptr()
        b.grad_perm = self.grad_perm.data_ptr()
        b.rev_flags = self.rev_flags.data_ptr()
        L.check(self.ctx.lib.pol_adjoint_bind(self.ctx.handle, C.byref(b)), self.ctx.handle, "pol_adjoint_bind")
        nseg = steps // checkpoint
        sd = cfg.storage_dtype
        fshape = tuple(self.ctx.f[0].shape)
        self.ckpt = None if self.store_all else torch.empty((nseg,) + fshape, dtype=sd, device=dev)
        self.seg = torch.empty((checkpoint,) + fshape, dtype=sd, device=dev)
        self._seg_holds = None  # the segment whose states self.seg holds (the forward run leaves the last one)
        self.timings = {}

    # ---- parameters -------------------------------------------------------------
    def set_permeability(self, K) -> None:
        self.ctx.perm.copy_(torch.as_tensor(K, dtype=self.ctx.config.real_dtype))

    def set_inlet_velocity(self, u: Sequence[float]) -> None:
        self.ctx.set_params(inlet_u=tuple(float(v) for v in u))

    # ---- forward ------------------------------------------------------------------
    def objective_step(self, n: int) -> bool:
        return n >= self.steps - self.window

    def forward(self, checkpoints: bool = False) -> float:
        """J (FP64) of one full run from rest; optionally store the checkpoints
        (or, with checkpoint=None, every state: the run then steps from one
        stored state into the next).  With checkpoints every K steps the run
        steps through the segment store over its last segment, so the reverse
        sweep starts without rebuilding it."""
        ctx = self.ctx
        ctx.set_current(0)
        ctx.init_equilibrium()
        J = 0.0
        self._seg_holds = None
        if checkpoints and self.store_all:
            seg = self.seg
            seg[0].copy_(ctx.f_current)
            for n in range(self.steps):
                out = seg[n + 1] if n + 1 < self.steps else ctx.f_next
                what = L.POL_WANT_REDUCE if self.objective_step(n) else 0
                red = ctx.step_from(seg[n], out, what)
                if what:
                    J += float(red[7])
            ctx.set_current(1 - ctx.current)  # the final state is the context's current one
            return J
        K = self.checkpoint
        last = self.steps - K  # first step of the last segment
        for n in range(self.steps):
            if checkpoints and n % K == 0:
                self.ckpt[n // K].copy_(ctx.f_current)
            if checkpoints and n >= last:
                # f^n .. f^{steps-1} into the segment store, the final state
                # into the context's next buffer (the same fused step as
                # ctx.step, from and into other buffers)
                seg = self.seg
                if n == last:
                    seg[0].copy_(ctx.f_current)
                out = seg[n - last + 1] if n + 1 < self.steps else ctx.f_next
                what = L.POL_WANT_REDUCE if self.objective_step(n) else 0
                red = ctx.step_from(seg[n - last], out, what)
                if what:
                    J += float(red[7])
                if n + 1 == self.steps:
                    self._seg_holds = self.steps // K - 1
                    ctx.set_current(1 - ctx.current)  # the final state is the context's current one
                continue
            if self.objective_step(n):
                red = ctx.step(1, L.POL_WANT_REDUCE)
                J += float(red[7])
            else:
                ctx.step(1)
        return J

    def forward_time(self) -> float:
        """Seconds for the plain forward run (no reductions, no checkpoints)."""
        ctx = self.ctx
        ctx.set_current(0)
        ctx.init_equilibrium()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.step(self.steps)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    # ---- reverse ------------------------------------------------------------------
    def gradient(self):
        """(J, dJ/dK [nz,ny,nx] float64 tensor, dJ/du_in [3]) by checkpointed adjoint."""
        ctx, lib = self.ctx, self.ctx.lib
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J = self.forward(checkpoints=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        L.check(lib.pol_adjoint_reset(ctx.handle, L.stream_handle()), ctx.handle, "pol_adjoint_reset")
        self.reverse_sweep()
        out = np.zeros(4, dtype=np.float64)
        L.check(lib.pol_adjoint_collect(ctx.handle, out.ctypes.data_as(C.c_void_p), L.stream_handle()), ctx.handle,
                "pol_adjoint_collect")
        t2 = time.perf_counter()
        self.timings = {"forward_s": t1 - t0, "reverse_s": t2 - t1, "J_adjoint": float(out[3])}
        return J, self.grad_perm.clone(), out[:3].copy()


    def _rebuild(self, s: int, seg: torch.Tensor) -> None:
        """f^{sK} .. f^{sK+K-1} into seg from checkpoint s (current stream), the
        K - 1 steps in one native call."""
        seg[0].copy_(self.ckpt[s])
        ctx = self.ctx
        stride = seg[0].numel() * seg.element_size()
        L.check(ctx.lib.pol_step_chain(ctx.handle, L.ptr(seg), stride, self.checkpoint, L.stream_handle()),
                ctx.handle, "pol_step_chain")

    def _sweep_segment(self, s: int, seg: torch.Tensor) -> None:
        """The K reverse steps of segment s in one native call."""
        ctx, K = self.ctx, self.checkpoint
        first = max(0, min(K, (self.steps - self.window) - s * K))  # first objective step within the segment
        stride = seg[0].numel() * seg.element_size()
        L.check(ctx.lib.pol_adjoint_sweep(ctx.handle, L.ptr(seg), stride, K, first, L.stream_handle()), ctx.handle,
                "pol_adjoint_sweep")

    def reverse_sweep(self) -> None:
        """Adjoint steps from the last step to the first (after forward with
        checkpoints and pol_adjoint_reset), enqueued on the current stream."""
        for s in reversed(range(self.steps // self.checkpoint)):
            if not self.store_all and self._seg_holds != s:
                self._rebuild(s, self.seg)  # f^{sK} .. f^{sK+K-1}
                self._seg_holds = s
            self._sweep_segment(s, self.seg)


class OutletFluxFunction(torch.autograd.Function):
    """J(K, u_in) as an autograd op; backward runs the checkpointed adjoint."""

    @staticmethod
    def forward(ctx, K: torch.Tensor, u_in: torch.Tensor, problem: AdjointProblem):
        problem.set_permeability(K.detach())
        problem.set_inlet_velocity(u_in.detach().double().cpu().tolist())
        J, gK, gu = problem.gradient()
        ctx.save_for_backward(gK.to(device=K.device, dtype=K.dtype),
                              torch.as_tensor(gu, dtype=u_in.dtype, device=u_in.device))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        gK, gu = ctx.saved_tensors
        g = grad_out.to(gK.device)
        return g.to(gK.dtype) * gK, g.to(gu.dtype) * gu, None


def outlet_flux(K: torch.Tensor, u_in: torch.Tensor, problem: AdjointProblem) -> torch.Tensor:
    return OutletFluxFunction.apply(K, u_in, problem)
