// Fused D3Q19 stream-collide kernel (pull scheme, SoA populations) for gfx950.
//
// One thread = one cell.  The thread gathers its 19 incoming populations
// f_q(x) = f*_q(x - e_q) from the previous post-collision buffer (halfway
// bounce-back f_q(x) = f*_{opp q}(x) on links whose source is a wall),
// computes moments, collides (BGK + Guo forcing + Smagorinsky + Darcy-
// Forchheimer, see cell_model.h) and writes its 19 post-collision values to
// the other buffer at its own index.  Every population is read once and
// written once: 152 B per fluid cell update in FP32, 76 B in 16-bit storage.
// Solid cells get no threads, except the solid tails of launched row chunks,
// which store zero deviations so that written lines are whole.
//
// Replaces (reference): cpu_backend.py:79-166 (collision + pull streaming),
// legacy/lbm_solver.py:488-654 (moments, collision, push streaming with
// halfway bounce-back, swap), ultra_optimized_lbm.py:383-525 (SoA fused form),
// les_turbulence.py:318-380 (eddy viscosity), filter_paper.py:471-536
// (porous resistance), visualizer.py:120-167 and numerical_stability.py:52-110
// (reductions, fused here as an FP64 wave64 epilogue).
#include "stream_collide.h"

namespace pol {

__global__ void __launch_bounds__(256) k_reduce_final(const double* __restrict__ partial, uint32_t nblk,
                                                      double* __restrict__ out) {
  __shared__ double sh[4][POL_NRED];
  double v[POL_NRED];
#pragma unroll
  for (int s = 0; s < POL_NRED; ++s) v[s] = red_identity(s);
  for (uint32_t b = threadIdx.x; b < nblk; b += 256) {
#pragma unroll
    for (int s = 0; s < POL_NRED; ++s) {
      const double p = partial[(size_t)b * POL_NRED + s];
      v[s] = red_is_max(s) ? fmax(v[s], p) : (red_is_min(s) ? fmin(v[s], p) : v[s] + p);
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int s = 0; s < POL_NRED; ++s) {
    const double r = red_is_max(s) ? wave_max(v[s]) : (red_is_min(s) ? wave_min(v[s]) : wave_sum(v[s]));
    if (lane == 0) sh[wid][s] = r;
  }
  __syncthreads();
  if (threadIdx.x < POL_NRED) {
    const int s = threadIdx.x;
    double r = sh[0][s];
    for (int w = 1; w < 4; ++w) r = red_is_max(s) ? fmax(r, sh[w][s]) : (red_is_min(s) ? fmin(r, sh[w][s]) : r + sh[w][s]);
    out[s] = r;
  }
}

template <typename S>
__global__ void __launch_bounds__(256) k_init_equilibrium(S* __restrict__ f, const typename Codec<S>::real* rho,
                                                          const typename Codec<S>::real* u, uint32_t N) {
  using C = Codec<S>;
  using R = typename C::real;
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= N) return;
  const R r = rho ? rho[i] : R(1);
  const R ux = u ? u[i] : R(0), uy = u ? u[(size_t)N + i] : R(0), uz = u ? u[2 * (size_t)N + i] : R(0);
  const R usq = ux * ux + uy * uy + uz * uz;
  static_for<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    C::store(f + (size_t)q * N + i, deq_q<R, q>(CW<R>(q), r - R(1), r, ux, uy, uz, usq));
  });
}

// Compact layout: every stored slot of the launched chunks takes feq of its
// dense cell's rho/u.
template <typename S>
__global__ void __launch_bounds__(BLOCK) k_init_equilibrium_sparse(const StepArgs a,
                                                                   const typename Codec<S>::real* rho,
                                                                   const typename Codec<S>::real* u) {
  using C = Codec<S>;
  using R = typename C::real;
  const uint32_t lblk = ordered_block(a.order, a.per_xcd);
  if (lblk >= a.nblk) return;
  uint32_t cell, x, y, z;
  int rb[9];
  int sx = 0;
  bool in_run = false, plain = false;
  const bool ok = item_record(lblk * BLOCK + threadIdx.x, a.n_items, a.chunks, a.records, a.cshift, a.N, a.dnx,
                              a.nx, cell, y, z, x, rb, sx, in_run, plain);
  if (!in_run) return;
  const uint32_t N = a.N, P = a.P;
  const uint32_t own = slot_clamp(rb[4] + sx, P);
  // padding slots past the row end take the rest state
  const R r = (rho && ok) ? rho[cell] : R(1);
  const R ux = (u && ok) ? u[cell] : R(0), uy = (u && ok) ? u[(size_t)N + cell] : R(0),
          uz = (u && ok) ? u[2 * (size_t)N + cell] : R(0);
  const R usq = ux * ux + uy * uy + uz * uz;
  S* __restrict__ f = (S*)a.fout;
  static_for<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    C::store(f + (size_t)q * P + own, deq_q<R, q>(CW<R>(q), r - R(1), r, ux, uy, uz, usq));
  });
}

template <typename S, bool A32>
static void launch_sc(const StepArgs& a, bool store_f, bool flags, bool full, dim3 grid, hipStream_t s) {
  dim3 block(BLOCK);
  if (store_f) {
    // streaming stores only in the 32-bit-offset kernels (the output is then
    // below 4 GiB but far above the last-level cache)
    const bool nt = A32 && a.stream_store;
    if (flags) {
      if (nt)
        hipLaunchKernelGGL((k_stream_collide<S, true, true, true, true, true>), grid, block, 0, s, a);
      else
        hipLaunchKernelGGL((k_stream_collide<S, true, true, true, A32>), grid, block, 0, s, a);
    } else if (full || !A32) {
      if (nt)
        hipLaunchKernelGGL((k_stream_collide<S, false, true, true, true, true>), grid, block, 0, s, a);
      else
        hipLaunchKernelGGL((k_stream_collide<S, false, true, true, A32>), grid, block, 0, s, a);
    } else {
      if (nt)
        hipLaunchKernelGGL((k_stream_collide<S, false, false, true, true, true>), grid, block, 0, s, a);
      else
        hipLaunchKernelGGL((k_stream_collide<S, false, false, true, true>), grid, block, 0, s, a);
    }
  } else {
    if (flags)
      hipLaunchKernelGGL((k_stream_collide<S, true, true, false, A32>), grid, block, 0, s, a);
    else if (full || !A32)
      hipLaunchKernelGGL((k_stream_collide<S, false, true, false, A32>), grid, block, 0, s, a);
    else
      hipLaunchKernelGGL((k_stream_collide<S, false, false, false, true>), grid, block, 0, s, a);
  }
}

template <typename S>
static hipError_t dispatch_sc(const StepArgs& a, bool store_f, const Grid& g, hipStream_t s) {
  const bool flags = a.flags != nullptr;
  const bool full =
      flags || (a.options & (POL_LES | POL_FORCE_FIELD | POL_PERM_FIELD | POL_GRAVITY | POL_PARTICLE_FIELD));
  // 32-bit offsets when a population buffer is below 4 GiB; larger grids take
  // the general (FULL) kernel with 64-bit addressing
  if (fits_addr32(Q, a.P, sizeof(S)))
    launch_sc<S, true>(a, store_f, flags, full, dim3(launch_blocks(a.order, g)), s);
  else
    launch_sc<S, false>(a, store_f, flags, full, dim3(launch_blocks(a.order, g)), s);
  return hipGetLastError();
}

hipError_t launch_stream_collide(int storage, const StepArgs& a, bool store_f, const Grid& g, hipStream_t s) {
  if (a.records) return launch_stream_collide_sparse(storage, a, store_f, g, s);
  switch (storage) {
    case POL_F32: return dispatch_sc<float>(a, store_f, g, s);
    case POL_F16: return dispatch_sc<_Float16>(a, store_f, g, s);
    case POL_BF16: return dispatch_sc<bf16_t>(a, store_f, g, s);
    case POL_F64: return dispatch_sc<double>(a, store_f, g, s);
  }
  return hipErrorInvalidValue;
}

template <typename S, bool A32>
static void launch_so(const StepArgs& a, bool flags, dim3 grid, hipStream_t s) {
  dim3 block(BLOCK);
  const bool nt = A32 && a.stream_store;
  if (flags) {
    if (nt)
      hipLaunchKernelGGL((k_stream_collide<S, true, true, true, true, true, false, true, false>), grid, block, 0, s, a);
    else
      hipLaunchKernelGGL((k_stream_collide<S, true, true, true, A32, false, false, true, false>), grid, block, 0, s, a);
  } else if (!A32) {
    hipLaunchKernelGGL((k_stream_collide<S, false, true, true, false, false, false, true, false>), grid, block, 0, s, a);
  } else if (nt) {
    hipLaunchKernelGGL((k_stream_collide<S, false, false, true, true, true, false, true, false>), grid, block, 0, s, a);
  } else {
    hipLaunchKernelGGL((k_stream_collide<S, false, false, true, true, false, false, true, false>), grid, block, 0, s, a);
  }
}

template <typename S>
static hipError_t dispatch_so(const StepArgs& a, const Grid& g, hipStream_t s) {
  const dim3 grid(launch_blocks(a.order, g));
  if (fits_addr32(Q, a.P, sizeof(S)))
    launch_so<S, true>(a, a.flags != nullptr, grid, s);
  else
    launch_so<S, false>(a, a.flags != nullptr, grid, s);
  return hipGetLastError();
}

hipError_t launch_stream_only(int storage, const StepArgs& a, const Grid& g, hipStream_t s) {
  if (a.records) return launch_stream_only_sparse(storage, a, g, s);
  switch (storage) {
    case POL_F32: return dispatch_so<float>(a, g, s);
    case POL_F16: return dispatch_so<_Float16>(a, g, s);
    case POL_BF16: return dispatch_so<bf16_t>(a, g, s);
    case POL_F64: return dispatch_so<double>(a, g, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_reduce_final(const double* partial, uint32_t nblk, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, s, partial, nblk, out);
  return hipGetLastError();
}

template <typename S>
static void launch_init(const StepArgs& a, const void* rho, const void* u, const Grid& g, hipStream_t s) {
  using R = typename Codec<S>::real;
  if (a.records) {
    hipLaunchKernelGGL(k_init_equilibrium_sparse<S>, dim3(launch_blocks(a.order, g)), dim3(BLOCK), 0, s, a, (const R*)rho,
                       (const R*)u);
  } else {
    hipLaunchKernelGGL(k_init_equilibrium<S>, dim3((a.N + 255) / 256), dim3(256), 0, s, (S*)a.fout, (const R*)rho,
                       (const R*)u, a.N);
  }
}

hipError_t launch_init_equilibrium(int storage, const StepArgs& a, const void* rho, const void* u, const Grid& g,
                                   hipStream_t s) {
  switch (storage) {
    case POL_F32: launch_init<float>(a, rho, u, g, s); break;
    case POL_F16: launch_init<_Float16>(a, rho, u, g, s); break;
    case POL_BF16: launch_init<bf16_t>(a, rho, u, g, s); break;
    case POL_F64: launch_init<double>(a, rho, u, g, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace pol
