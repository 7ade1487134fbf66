// Fused D3Q19 stream-collide kernel (pull scheme, SoA populations) for gfx950.
//
// One thread = one cell.  The thread gathers its 19 incoming populations
// f_q(x) = f*_q(x - e_q) from the previous post-collision buffer (halfway
// bounce-back f_q(x) = f*_{opp q}(x) on links whose source is a wall),
// computes moments, collides (BGK + Guo forcing + Smagorinsky + Darcy-
// Forchheimer, see cell_model.h) and writes its 19 post-collision values to
// the other buffer at its own index.  Every population is read once and
// written once: 152 B per fluid cell update in FP32, 76 B in 16-bit storage;
// solid cells are never read or written.
//
// Replaces (reference): cpu_backend.py:79-166 (collision + pull streaming),
// legacy/lbm_solver.py:488-654 (moments, collision, push streaming with
// halfway bounce-back, swap), ultra_optimized_lbm.py:383-525 (SoA fused form),
// les_turbulence.py:318-380 (eddy viscosity), filter_paper.py:471-536
// (porous resistance), visualizer.py:120-167 and numerical_stability.py:52-110
// (reductions, fused here as an FP64 wave64 epilogue).
#include "cell_model.h"
#include "engine.h"

namespace pol {

template <typename S, bool FLAGS, bool FULL, bool STORE, bool A32, bool NT = false>
__global__ void __launch_bounds__(BLOCK) k_stream_collide(const StepArgs a) {
  using C = Codec<S>;
  using R = typename C::real;
  using AD = Addr<A32>;
  using AI = typename AD::idx;
  const uint32_t lblk = logical_block(a.per_xcd);
  if (lblk >= a.nblk) return;  // whole padding block exits (uniform)
  uint32_t cell, yz, x;
  const bool inside = item_cell(lblk * BLOCK + threadIdx.x, a.n_items, a.chunks, a.cshift, a.N, a.dnx, a.nx, cell, yz, x);
  const uint32_t z = fdiv(yz, a.dny);
  const uint32_t y = yz - z * a.ny;

  const uint32_t fl = FLAGS ? (inside ? a.flags[cell] : (POL_KIND_SOLID << POL_KIND_SHIFT)) : 0u;
  const uint32_t kind = (fl >> POL_KIND_SHIFT) & POL_KIND_MASK;
  const bool active = inside && kind != POL_KIND_SOLID;

  Phys<R> P;
  P.tau0 = (R)a.tau0;
  P.nu0 = (R)a.nu0;
  P.les_c = (R)a.les_c;
  P.tau_max = (R)a.tau_max;
  P.G[0] = (R)a.G[0];
  P.G[1] = (R)a.G[1];
  P.G[2] = (R)a.G[2];
  P.K = (R)a.K;
  P.CF = (R)a.CF;
  P.inlet_rho = (R)a.inlet_rho;
  P.inlet_u[0] = (R)a.inlet_u[0];
  P.inlet_u[1] = (R)a.inlet_u[1];
  P.inlet_u[2] = (R)a.inlet_u[2];
  P.outlet_rho = (R)a.outlet_rho;

  const S* __restrict__ fin = (const S*)a.fin;
  S* __restrict__ fout = (S*)a.fout;
  const uint32_t N = a.N;

  R f[Q];  // scaled deviations 36 (f - w)
  CellOut<R> o{R(0), R(0), R(0), R(0)};
  if (active) {
    if (!FULL || kind != POL_KIND_INLET) {
      // neighbour offsets; non-periodic axes clamp to the cell's own plane
      // (zero-gradient for open-boundary cells; walls use bounce-back bits)
      const bool px = a.options & POL_PERIODIC_X, py = a.options & POL_PERIODIC_Y, pz = a.options & POL_PERIODIC_Z;
      const int nx = (int)a.nx, nxy = (int)a.nxy;
      const int oxm = x == 0 ? (px ? nx - 1 : 0) : -1;               // source at x-1
      const int oxp = x == a.nx - 1 ? (px ? -(nx - 1) : 0) : 1;      // source at x+1
      const int oym = y == 0 ? (py ? ((int)a.ny - 1) * nx : 0) : -nx;
      const int oyp = y == a.ny - 1 ? (py ? -((int)a.ny - 1) * nx : 0) : nx;
      const int ozm = z == 0 ? (pz ? ((int)a.nz - 1) * nxy : 0) : -nxy;
      const int ozp = z == a.nz - 1 ? (pz ? -((int)a.nz - 1) * nxy : 0) : nxy;
      static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int off = (EX(q) > 0 ? oxm : (EX(q) < 0 ? oxp : 0)) + (EY(q) > 0 ? oym : (EY(q) < 0 ? oyp : 0)) +
                        (EZ(q) > 0 ? ozm : (EZ(q) < 0 ? ozp : 0));
        AI src = (AI)q * N + (uint32_t)((int)cell + off);
        if constexpr (FLAGS && q != 0) {
          if (fl & (1u << q)) src = (AI)OPP(q) * N + cell;
        }
        // bounce-back reads D*_{opp q}(x); w_{opp q} = w_q, so the stored
        // deviation of direction opp(q) is the deviation of link q.
        f[q] = C::load(AD::at(fin, src));
      });
    }
    R fx = R(0), fy = R(0), fz = R(0), K = P.K;
    R bx = R(0), by = R(0), bz = R(0), B = R(0);
    if constexpr (FULL) {
      if (a.options & POL_FORCE_FIELD) {
        const R* F = (const R*)a.force;
        fx = F[cell];
        fy = F[(size_t)N + cell];
        fz = F[2 * (size_t)N + cell];
      }
      if (a.options & POL_PARTICLE_FIELD) {
        const R* Pd = (const R*)a.pdrag;
        bx = Pd[cell];
        by = Pd[(size_t)N + cell];
        bz = Pd[2 * (size_t)N + cell];
        B = Pd[3 * (size_t)N + cell];
      }
    }
    const bool porous = FULL && (fl & POL_FLAG_POROUS);
    if constexpr (FULL) {
      if (porous && (a.options & POL_PERM_FIELD)) K = ((const R*)a.perm)[cell];
    }
    const bool les = FULL && (a.options & POL_LES) && !(fl & POL_FLAG_NO_LES);
    if constexpr (STORE) {
      o = collide_cell<R, FULL>(f, kind, porous, les, K, fx, fy, fz, bx, by, bz, B, P);
    } else {
      // moments only: the same velocity definition the collision uses
      o = collide_cell<R, FULL>(f, kind, porous, false, K, fx, fy, fz, bx, by, bz, B, P);
    }
  }
  if constexpr (STORE) {
    // Solid lanes next to fluid ones store zero deviations (f = w; no gather
    // ever reads a solid cell), so that every 128-byte line of the output is
    // written whole: a partly written line costs HBM a read-modify-write.
    // Chunk launches list only chunks holding fluid, so all their lanes store;
    // dense launches store a solid lane when its 32-lane group holds fluid.
    bool store = active;
    if constexpr (FLAGS) {
      const uint64_t any = __ballot(active);
      const uint64_t group = 0xffffffffull << (threadIdx.x & 32);
      store = inside && (a.chunks != nullptr || (any & group) != 0);
      if (!active) {
#pragma unroll
        for (int q = 0; q < Q; ++q) f[q] = R(0);
      }
    }
    if (store) {
      static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (NT) C::store_nt(AD::at(fout, (AI)q * N + cell), f[q]);
        else C::store(AD::at(fout, (AI)q * N + cell), f[q]);
      });
    }
  }

  if (a.what & POL_WANT_MACRO) {
    if (inside) {
      R* rho = (R*)a.rho;
      R* u = (R*)a.u;
      rho[cell] = R(1) + o.rd;
      u[cell] = o.ux;
      u[(size_t)N + cell] = o.uy;
      u[2 * (size_t)N + cell] = o.uz;
    }
  }
  if (a.what & POL_WANT_REDUCE) {
    double v[POL_NRED];
#pragma unroll
    for (int s = 0; s < POL_NRED; ++s) v[s] = red_identity(s);
    if (active) {
      const double rho = 1.0 + (double)o.rd, ux = (double)o.ux, uy = (double)o.uy, uz = (double)o.uz;
      const double usq = ux * ux + uy * uy + uz * uz;
      const bool finite = isfinite(rho) && isfinite(usq);
      v[POL_RED_MASS] = rho;
      v[POL_RED_MOMX] = rho * ux;
      v[POL_RED_MOMY] = rho * uy;
      v[POL_RED_MOMZ] = rho * uz;
      v[POL_RED_KE] = 0.5 * rho * usq;
      v[POL_RED_UMAX] = finite ? sqrt(usq) : -1.0e300;
      v[POL_RED_NONFINITE] = finite ? 0.0 : 1.0;
      v[POL_RED_OUTFLUX] = kind == POL_KIND_OUTLET ? -rho * uz : 0.0;
      v[POL_RED_CELLS] = 1.0;
      v[POL_RED_RHOMIN] = finite ? rho : 1.0e300;
      v[POL_RED_RHOMAX] = finite ? rho : -1.0e300;
      v[POL_RED_INFLUX] = kind == POL_KIND_INLET ? -rho * uz : 0.0;
    }
    block_reduce_store<BLOCK / 64>(v, a.partial, lblk);
  }
}

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
  if (fits_addr32(Q, a.N, sizeof(S)))
    launch_sc<S, true>(a, store_f, flags, full, dim3(g.launch), s);
  else
    launch_sc<S, false>(a, store_f, flags, full, dim3(g.launch), s);
  return hipGetLastError();
}

hipError_t launch_stream_collide(int storage, const StepArgs& a, bool store_f, const Grid& g, hipStream_t s) {
  switch (storage) {
    case POL_F32: return dispatch_sc<float>(a, store_f, g, s);
    case POL_F16: return dispatch_sc<_Float16>(a, store_f, g, s);
    case POL_BF16: return dispatch_sc<bf16_t>(a, store_f, g, s);
    case POL_F64: return dispatch_sc<double>(a, store_f, g, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_reduce_final(const double* partial, uint32_t nblk, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, s, partial, nblk, out);
  return hipGetLastError();
}

hipError_t launch_init_equilibrium(int storage, void* f, const void* rho, const void* u, uint32_t N, hipStream_t s) {
  dim3 grid((N + 255) / 256), block(256);
  switch (storage) {
    case POL_F32:
      hipLaunchKernelGGL(k_init_equilibrium<float>, grid, block, 0, s, (float*)f, (const float*)rho, (const float*)u, N);
      break;
    case POL_F16:
      hipLaunchKernelGGL(k_init_equilibrium<_Float16>, grid, block, 0, s, (_Float16*)f, (const float*)rho,
                         (const float*)u, N);
      break;
    case POL_BF16:
      hipLaunchKernelGGL(k_init_equilibrium<bf16_t>, grid, block, 0, s, (bf16_t*)f, (const float*)rho,
                         (const float*)u, N);
      break;
    case POL_F64:
      hipLaunchKernelGGL(k_init_equilibrium<double>, grid, block, 0, s, (double*)f, (const double*)rho,
                         (const double*)u, N);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace pol
