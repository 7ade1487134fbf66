// Field diagnostics of the solver state as fused device reductions.
//
// Replaces (reference): the host-side numpy / Taichi reductions of
//   numerical_stability.py:53-110   (NaN / Inf counts, rho range, max |u|
//                                    over fluid cells, populations included)
//   numerical_stability.py:165-197  (emergency stabilisation)
//   lbm_diagnostics.py:151-460      (Mach / density quality, conservation,
//                                    regional |u|, outlet flow, z profile,
//                                    temporal change)
//   visualizer.py:120-167           (interior water / air mass, |u| range)
//   enhanced_visualizer.py:1487-1505 (|u| and pressure fields of the export)
// which copy rho, u (and f) to the host, about 180 MB per diagnostic step at
// 224^3.  Here one pass reads each field once on the device.
//
// Statistics kernel: a grid of (tiles, nz) blocks, block (t, z) looping over
// plane z; per-lane FP64 accumulators, wave64 shuffles, one partial per block
// per statistic, folded in block order by one block (deterministic).  The
// per-plane density sums of the z profile fall out of the same partials.
// |u| is formed in FP32 as sqrt((ux*ux + uy*uy) + uz*uz) with explicit
// round-to-nearest operations (no contraction), the value the reference's
// float32 numpy expressions give per element.
#include "engine.h"
#include "cell_model.h"

namespace pol {

enum StatKind { SK_SUM = 0, SK_MIN = 1, SK_MAX = 2, SK_NANMAX = 3 };

__host__ __device__ constexpr int stat_kind(int s) {
  return (s == POL_STAT_MAX_ABS_DRHO || s == POL_STAT_MAX_UMAG)                                  ? SK_NANMAX
         : (s == POL_STAT_FLUID_RHO_MIN || s == POL_STAT_INT_UMIN)                               ? SK_MIN
         : (s == POL_STAT_FLUID_RHO_MAX || s == POL_STAT_FLUID_UMAX || s == POL_STAT_INT_UMAX) ? SK_MAX
                                                                                               : SK_SUM;
}
__host__ __device__ constexpr double stat_identity(int s) {
  return s == POL_STAT_FLUID_UMAX ? 0.0
         : stat_kind(s) == SK_MIN ? 1.0e300
         : (stat_kind(s) == SK_MAX || stat_kind(s) == SK_NANMAX) ? -1.0e300
                                                                 : 0.0;
}
__device__ inline double nanmax(double a, double b) { return (isnan(a) || isnan(b)) ? NAN : fmax(a, b); }
__device__ inline double stat_combine(int s, double a, double b) {
  switch (stat_kind(s)) {
    case SK_MIN: return fmin(a, b);
    case SK_MAX: return fmax(a, b);
    case SK_NANMAX: return nanmax(a, b);
    default: return a + b;
  }
}
template <int S>
__device__ inline double combine_c(double a, double b) {
  if constexpr (stat_kind(S) == SK_MIN) return fmin(a, b);
  else if constexpr (stat_kind(S) == SK_MAX) return fmax(a, b);
  else if constexpr (stat_kind(S) == SK_NANMAX) return nanmax(a, b);
  else return a + b;
}
__device__ inline double wave_combine(int s, double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = stat_combine(s, v, __shfl_xor(v, o, 64));
  return v;
}

// |u| in FP32 without contraction: ((ux*ux) + (uy*uy)) + (uz*uz), each
// operation rounded on its own (no FMA), then the correctly rounded sqrt
__device__ inline float umag_f32(float ux, float uy, float uz) {
#pragma clang fp contract(off)
  const float s = (ux * ux + uy * uy) + uz * uz;
  return __builtin_sqrtf(s);
}
__device__ inline double umag_f32(double ux, double uy, double uz) {
  return umag_f32((float)ux, (float)uy, (float)uz);
}

// Count statistics accumulate as 32-bit integers (fewer registers than FP64).
__host__ __device__ constexpr bool stat_is_count(int s) {
  return s == POL_STAT_CELLS || s == POL_STAT_ACTIVE || s == POL_STAT_MACH_GT || s == POL_STAT_FLUID ||
         s == POL_STAT_FLUID_RHO_NAN || s == POL_STAT_FLUID_RHO_INF || s == POL_STAT_FLUID_U_NAN ||
         s == POL_STAT_FLUID_U_INF || s == POL_STAT_INT_CELLS || s == POL_STAT_UMAG_ACTIVE;
}

template <typename R>
struct CellIn {
  R r, ux, uy, uz, pr, pux, puy, puz, ph;
  uint32_t fl;
};

template <typename R>
__device__ inline void load_cell(const StatsArgs& a, size_t cell, CellIn<R>& c) {
  const R* rho = (const R*)a.rho;
  const R* u = (const R*)a.u;
  const size_t N = a.N;
  c.r = rho[cell];
  c.ux = u[cell];
  c.uy = u[N + cell];
  c.uz = u[2 * N + cell];
  c.fl = a.flags ? a.flags[cell] : (uint32_t)(POL_KIND_FLUID << POL_KIND_SHIFT);
  if (a.prev_rho) c.pr = ((const R*)a.prev_rho)[cell];
  if (a.prev_u) {
    const R* pu = (const R*)a.prev_u;
    c.pux = pu[cell];
    c.puy = pu[N + cell];
    c.puz = pu[2 * N + cell];
  }
  if (a.phase) c.ph = ((const R*)a.phase)[cell];
}

template <typename R>
__global__ void __launch_bounds__(BLOCK) k_field_stats(const StatsArgs a) {
  const uint32_t z = blockIdx.y;
  double v[POL_NSTAT];
  uint32_t cnt[POL_NSTAT];
#pragma unroll
  for (int s = 0; s < POL_NSTAT; ++s) {
    v[s] = stat_identity(s);
    cnt[s] = 0u;
  }
  double zs = 0.0;
  const float cs = (float)0.57735026918962573;  // float32(1 / sqrt(3)), the reference's c_s
  const uint32_t stride = a.tiles * BLOCK;
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  CellIn<R> c{}, nxt{};
  if (i < a.nxy) load_cell<R>(a, (size_t)z * a.nxy + i, c);
  // software-pipelined: the next cell's loads are in flight while this one reduces
  for (; i < a.nxy; i += stride) {
    const uint32_t in = i + stride;
    if (in < a.nxy) load_cell<R>(a, (size_t)z * a.nxy + in, nxt);
    const uint32_t y = i / a.nx, x = i - y * a.nx;
    const double rd = (double)c.r;
    const double um = (double)umag_f32(c.ux, c.uy, c.uz);
    cnt[POL_STAT_CELLS] += 1u;
    const double d = rd - 1.0;
    v[POL_STAT_SUM_DRHO] += d;
    v[POL_STAT_SUM_DRHO2] += d * d;
    v[POL_STAT_MAX_ABS_DRHO] = nanmax(v[POL_STAT_MAX_ABS_DRHO], fabs(d));
    v[POL_STAT_MOM_X] += rd * (double)c.ux;
    v[POL_STAT_MOM_Y] += rd * (double)c.uy;
    v[POL_STAT_MOM_Z] += rd * (double)c.uz;
    v[POL_STAT_SUM_UMAG] += um;
    v[POL_STAT_SUM_UMAG2] += um * um;
    v[POL_STAT_MAX_UMAG] = nanmax(v[POL_STAT_MAX_UMAG], um);
    const float mach = (float)um / cs;
    if (mach > 1e-10f) {
      v[POL_STAT_SUM_MACH_ACTIVE] += (double)mach;
      cnt[POL_STAT_ACTIVE] += 1u;
    }
    if (mach > 0.1f) cnt[POL_STAT_MACH_GT] += 1u;
    if ((float)um > 1e-10f) {
      v[POL_STAT_SUM_UMAG_ACTIVE] += um;
      cnt[POL_STAT_UMAG_ACTIVE] += 1u;
    }
    v[POL_STAT_SUM_UX] += (double)c.ux;
    v[POL_STAT_SUM_UY] += (double)c.uy;
    v[POL_STAT_SUM_UZ] += (double)c.uz;
    v[POL_STAT_SUM_UX2] += (double)c.ux * (double)c.ux;
    v[POL_STAT_SUM_UY2] += (double)c.uy * (double)c.uy;
    v[POL_STAT_SUM_UZ2] += (double)c.uz * (double)c.uz;
    if ((int)z >= a.z_top) v[POL_STAT_SUM_UMAG_TOP] += um;
    else if ((int)z >= a.z_mid) v[POL_STAT_SUM_UMAG_MID] += um;
    else v[POL_STAT_SUM_UMAG_BOT] += um;
    if ((int)z < a.z_out) v[POL_STAT_OUTLET_DOWN] += fmax(-(double)c.uz, 0.0);
    if (a.prev_rho) {
      const double e = rd - (double)c.pr;
      v[POL_STAT_DRHO_PREV2] += e * e;
    }
    if (a.prev_u) {
      const double ex = (double)c.ux - (double)c.pux, ey = (double)c.uy - (double)c.puy,
                   ez = (double)c.uz - (double)c.puz;
      v[POL_STAT_DU_PREV2] += ex * ex + ey * ey + ez * ez;
    }
    zs += rd;
    const uint32_t kind = (c.fl >> POL_KIND_SHIFT) & POL_KIND_MASK;
    if (kind != POL_KIND_SOLID) {
      cnt[POL_STAT_FLUID] += 1u;
      if (isnan(rd)) cnt[POL_STAT_FLUID_RHO_NAN] += 1u;
      else if (isinf(rd)) cnt[POL_STAT_FLUID_RHO_INF] += 1u;
      else {
        v[POL_STAT_FLUID_RHO_MIN] = fmin(v[POL_STAT_FLUID_RHO_MIN], rd);
        v[POL_STAT_FLUID_RHO_MAX] = fmax(v[POL_STAT_FLUID_RHO_MAX], rd);
      }
      if (isnan(um)) cnt[POL_STAT_FLUID_U_NAN] += 1u;
      else if (isinf(um)) cnt[POL_STAT_FLUID_U_INF] += 1u;
      else v[POL_STAT_FLUID_UMAX] = fmax(v[POL_STAT_FLUID_UMAX], um);
      if (x >= 1 && x + 1 < a.nx && y >= 1 && y + 1 < a.ny && z >= 1 && z + 1 < a.nz) {
        cnt[POL_STAT_INT_CELLS] += 1u;
        const bool water = a.phase ? (c.ph > R(0.1)) : true;
        if (water) v[POL_STAT_INT_WATER] += rd;
        else v[POL_STAT_INT_AIR] += rd;
        v[POL_STAT_INT_UMAX] = fmax(v[POL_STAT_INT_UMAX], um);
        v[POL_STAT_INT_UMIN] = fmin(v[POL_STAT_INT_UMIN], um);
        v[POL_STAT_INT_USUM] += um;
      }
    }
    c = nxt;
  }
  static_for<0, POL_NSTAT>([&](auto sc) {
    constexpr int S = decltype(sc)::value;
    if constexpr (stat_is_count(S)) v[S] = (double)cnt[S];
  });
  // compile-time slot indices keep the accumulators in registers
  __shared__ double sh[BLOCK / 64][POL_NSTAT + 1];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  static_for<0, POL_NSTAT>([&](auto sc) {
    constexpr int S = decltype(sc)::value;
    double r = v[S];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) r = combine_c<S>(r, __shfl_xor(r, o, 64));
    if (lane == 0) sh[wid][S] = r;
  });
  zs = wave_sum(zs);
  if (lane == 0) sh[wid][POL_NSTAT] = zs;
  __syncthreads();
  const size_t blk = (size_t)z * a.tiles + blockIdx.x;
  if (threadIdx.x < POL_NSTAT) {
    const int s = threadIdx.x;
    double r = sh[0][s];
    for (int w = 1; w < BLOCK / 64; ++w) r = stat_combine(s, r, sh[w][s]);
    a.partial[blk * POL_NSTAT + s] = r;
  } else if (threadIdx.x == POL_NSTAT) {
    double r = sh[0][POL_NSTAT];
    for (int w = 1; w < BLOCK / 64; ++w) r += sh[w][POL_NSTAT];
    a.zpart[blk] = r;
  }
}

// Fold the block partials: block s < POL_NSTAT folds statistic s, block
// POL_NSTAT + z the tiles of plane z.  Thread t combines the partials
// t, t + 256, ... in order, then the block combines its threads in a fixed
// tree: the result does not depend on timing (deterministic).
__global__ void __launch_bounds__(256) k_stats_final(const double* __restrict__ partial,
                                                     const double* __restrict__ zpart, uint32_t nblk, uint32_t tiles,
                                                     double* __restrict__ out, double* __restrict__ zout) {
  const uint32_t b = blockIdx.x;
  const bool stat = b < POL_NSTAT;
  const int s = stat ? (int)b : 0;
  double r = stat ? stat_identity(s) : 0.0;
  if (stat) {
    for (uint32_t k = threadIdx.x; k < nblk; k += 256) r = stat_combine(s, r, partial[(size_t)k * POL_NSTAT + s]);
  } else {
    const uint32_t z = b - POL_NSTAT;
    for (uint32_t k = threadIdx.x; k < tiles; k += 256) r += zpart[(size_t)z * tiles + k];
  }
  r = stat ? wave_combine(s, r) : wave_sum(r);
  __shared__ double sh[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = sh[0];
    for (int w = 1; w < 4; ++w) t = stat ? stat_combine(s, t, sh[w]) : t + sh[w];
    if (stat) out[s] = t;
    else zout[b - POL_NSTAT] = t;
  }
}

// Work item -> (dense cell, population slot, non-solid) for the bound launch
// list; the compact layout's slot comes from the chunk record.
template <bool SPARSE>
__device__ inline bool work_cell(const StepArgs& a, uint32_t lblk, uint32_t& cell, uint32_t& own, uint32_t& fl) {
  uint32_t yz, x;
  int rb[9];
  int sx = 0;
  bool in_run = false, inside;
  if constexpr (SPARSE)
    inside = item_record(lblk * BLOCK + threadIdx.x, a.n_items, a.records, a.cshift, a.N, a.dnx, a.nx, cell, yz, x, rb,
                         sx, in_run);
  else
    inside = item_cell(lblk * BLOCK + threadIdx.x, a.n_items, a.chunks, a.cshift, a.N, a.dnx, a.nx, cell, yz, x);
  own = cell;
  if constexpr (SPARSE) own = slot_clamp(rb[4] + sx, a.P);
  fl = (inside && a.flags) ? a.flags[cell] : (inside ? (POL_KIND_FLUID << POL_KIND_SHIFT) : 0u);
  const uint32_t kind = (fl >> POL_KIND_SHIFT) & POL_KIND_MASK;
  return inside && kind != POL_KIND_SOLID;
}

// NaN / Inf populations of the current state over non-solid cells
// (numerical_stability.py:92-98); block partials {nan, inf}.
template <typename S, bool SPARSE>
__global__ void __launch_bounds__(BLOCK) k_pop_nonfinite(const StepArgs a, double* __restrict__ partial) {
  const uint32_t lblk = logical_block(a.per_xcd);
  if (lblk >= a.nblk) return;
  uint32_t cell, own, fl;
  double nn = 0.0, ni = 0.0;
  if (work_cell<SPARSE>(a, lblk, cell, own, fl)) {
    const S* __restrict__ f = (const S*)a.fin;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const double d = (double)Codec<S>::load(f + (size_t)q * a.P + own);
      nn += isnan(d) ? 1.0 : 0.0;
      ni += isinf(d) ? 1.0 : 0.0;
    }
  }
  __shared__ double sh[BLOCK / 64][2];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  nn = wave_sum(nn);
  ni = wave_sum(ni);
  if (lane == 0) {
    sh[wid][0] = nn;
    sh[wid][1] = ni;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double r = 0.0;
    for (int w = 0; w < BLOCK / 64; ++w) r += sh[w][threadIdx.x];
    partial[(size_t)lblk * 2 + threadIdx.x] = r;
  }
}

__global__ void __launch_bounds__(64) k_pop_final(const double* __restrict__ partial, uint32_t nblk,
                                                  double* __restrict__ out) {
  if (threadIdx.x < 2) {
    double r = 0.0;
    for (uint32_t b = 0; b < nblk; ++b) r += partial[(size_t)b * 2 + threadIdx.x];
    out[POL_STAT_FLUID_POP_NAN + threadIdx.x] = r;
  }
}

// Emergency stabilisation (numerical_stability.py:165-197) of non-solid cells.
template <typename S, bool SPARSE>
__global__ void __launch_bounds__(BLOCK) k_stabilize(const StepArgs a, float max_u, float min_rho, float max_rho) {
  using C = Codec<S>;
  using R = typename C::real;
  const uint32_t lblk = logical_block(a.per_xcd);
  if (lblk >= a.nblk) return;
  uint32_t cell, own, fl;
  if (!work_cell<SPARSE>(a, lblk, cell, own, fl)) return;
  R* __restrict__ rho = (R*)a.rho;
  R* __restrict__ u = (R*)a.u;
  const size_t N = a.N;
  R r = rho[cell];
  if (isnan(r) || isinf(r) || r <= R(0)) r = R(1);
  else if (r > (R)max_rho) r = (R)max_rho;
  else if (r < (R)min_rho) r = (R)min_rho;
  rho[cell] = r;
  R ux = u[cell], uy = u[N + cell], uz = u[2 * N + cell];
  const R um = (R)umag_f32(ux, uy, uz);
  if (isnan(um) || isinf(um)) {
    ux = uy = uz = R(0);
  } else if (um > (R)max_u) {
    const R s = (R)max_u / um;
    ux *= s;
    uy *= s;
    uz *= s;
  }
  u[cell] = ux;
  u[N + cell] = uy;
  u[2 * N + cell] = uz;
  // non-finite populations -> equilibrium of the repaired (rho, u)
  S* __restrict__ f = (S*)a.fout;
  const R usq = ux * ux + uy * uy + uz * uz;
  static_for<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    S* p = f + (size_t)q * a.P + own;
    const R d = C::load(p);
    if (isnan(d) || isinf(d)) C::store(p, deq_q<R, q>(CW<R>(q), r - R(1), r, ux, uy, uz, usq));
  });
}

// |u| and pressure fields of the export (enhanced_visualizer.py:1493-1505).
template <typename R>
__global__ void __launch_bounds__(256) k_export_fields(const R* __restrict__ rho, const R* __restrict__ u, uint32_t N,
                                                       R* __restrict__ mag, R* __restrict__ pres) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= N) return;
  const R ux = u[i], uy = u[(size_t)N + i], uz = u[2 * (size_t)N + i];
  if constexpr (sizeof(R) == 4) {
    mag[i] = umag_f32(ux, uy, uz);
    pres[i] = rho[i] * (float)(1.0 / 3.0);
  } else {
    mag[i] = sqrt((ux * ux + uy * uy) + uz * uz);
    pres[i] = rho[i] * (1.0 / 3.0);
  }
}

// ---- launchers ------------------------------------------------------------
hipError_t launch_field_stats(bool real_double, const StatsArgs& s, double* out_dev, double* zout_dev,
                              hipStream_t st) {
  const dim3 grid(s.tiles, s.nz);
  if (real_double)
    hipLaunchKernelGGL(k_field_stats<double>, grid, dim3(BLOCK), 0, st, s);
  else
    hipLaunchKernelGGL(k_field_stats<float>, grid, dim3(BLOCK), 0, st, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_stats_final, dim3(POL_NSTAT + s.nz), dim3(256), 0, st, s.partial, s.zpart, s.tiles * s.nz,
                     s.tiles, out_dev, zout_dev);
  return hipGetLastError();
}

template <typename S>
static void launch_pop(const StepArgs& a, double* partial, dim3 grid, hipStream_t st) {
  if (a.records)
    hipLaunchKernelGGL((k_pop_nonfinite<S, true>), grid, dim3(BLOCK), 0, st, a, partial);
  else
    hipLaunchKernelGGL((k_pop_nonfinite<S, false>), grid, dim3(BLOCK), 0, st, a, partial);
}

hipError_t launch_pop_nonfinite(int storage, const StepArgs& a, double* partial, double* out_dev, const Grid& g,
                                hipStream_t st) {
  const dim3 grid(g.launch);
  switch (storage) {
    case POL_F32: launch_pop<float>(a, partial, grid, st); break;
    case POL_F16: launch_pop<_Float16>(a, partial, grid, st); break;
    case POL_BF16: launch_pop<bf16_t>(a, partial, grid, st); break;
    case POL_F64: launch_pop<double>(a, partial, grid, st); break;
    default: return hipErrorInvalidValue;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_pop_final, dim3(1), dim3(64), 0, st, partial, g.nblk, out_dev);
  return hipGetLastError();
}

template <typename S>
static void launch_stab(const StepArgs& a, float mu, float rmin, float rmax, dim3 grid, hipStream_t st) {
  if (a.records)
    hipLaunchKernelGGL((k_stabilize<S, true>), grid, dim3(BLOCK), 0, st, a, mu, rmin, rmax);
  else
    hipLaunchKernelGGL((k_stabilize<S, false>), grid, dim3(BLOCK), 0, st, a, mu, rmin, rmax);
}

hipError_t launch_stabilize(int storage, const StepArgs& a, float max_u, float min_rho, float max_rho, const Grid& g,
                            hipStream_t st) {
  const dim3 grid(g.launch);
  switch (storage) {
    case POL_F32: launch_stab<float>(a, max_u, min_rho, max_rho, grid, st); break;
    case POL_F16: launch_stab<_Float16>(a, max_u, min_rho, max_rho, grid, st); break;
    case POL_BF16: launch_stab<bf16_t>(a, max_u, min_rho, max_rho, grid, st); break;
    case POL_F64: launch_stab<double>(a, max_u, min_rho, max_rho, grid, st); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_export_fields(bool real_double, const void* rho, const void* u, uint32_t N, void* mag, void* pres,
                                hipStream_t st) {
  const dim3 grid((N + 255) / 256);
  if (real_double)
    hipLaunchKernelGGL(k_export_fields<double>, grid, dim3(256), 0, st, (const double*)rho, (const double*)u, N,
                       (double*)mag, (double*)pres);
  else
    hipLaunchKernelGGL(k_export_fields<float>, grid, dim3(256), 0, st, (const float*)rho, (const float*)u, N,
                       (float*)mag, (float*)pres);
  return hipGetLastError();
}

}  // namespace pol
