// Thermal D3Q7 lattice Boltzmann step (temperature populations) for gfx950.
//
// Replaces (reference): src/physics/thermal_lbm.py
//   :186-201 equilibrium g_eq = w_q T; :203-229 BGK collision with the heat
//   source; :231-253 streaming (an out-of-box source keeps the cell's own
//   value); :255-266 temperature; :268-286 heat flux -k grad T; :288-346
//   Dirichlet / Neumann / convective boundary forms; :348-376 stability test;
//   :378-394 min / max / mean; :518-538 convection source -u.grad T.
//
// The reference collides locally and then pull-streams in a second pass.
// Here one kernel collides and pushes: cell x reads its 7 streamed
// populations (coalesced), forms T, S and g*, and writes g*_q to x + e_q, or
// to itself when x - e_q lies outside the box -- the reference's boundary
// rule, every destination written exactly once.  A second kernel forms the
// new temperature (the reference's compute_temperature) with the stability
// test fused in; in a run of pure-diffusion steps each later collide pass
// stores the temperature of its input state instead, and the second kernel
// runs once, after the last step.  Temperature buffers ping-pong, so
// temperature_old needs no copy.  The arithmetic follows the reference's FP32 expressions operation by
// operation with contraction off:
//   T = ((((((0 + g0) + g1) + g2) + g3) + g4) + g5) + g6
//   g* = (g - omega (g - w T)) + (w S) 1
//   dT/dx = (T[x+1] - T[x-1]) / 2,  S += -((ux dTdx + uy dTdy) + uz dTdz)
// so an FP32 numpy restatement of the same expressions agrees bit for bit.
// With a base heat source bound (hsb) every step starts from it, S = base
// (+ convection), and writes S to the heat source: the coupled loop's
// reset_heat_source_to_base (thermal_fluid_coupled.py:193-194) at no extra
// pass.
#include "engine.h"

namespace pol {

constexpr int TQ = 7;
__host__ __device__ constexpr int TEX(int q) { return q == 1 ? 1 : (q == 2 ? -1 : 0); }
__host__ __device__ constexpr int TEY(int q) { return q == 3 ? 1 : (q == 4 ? -1 : 0); }
__host__ __device__ constexpr int TEZ(int q) { return q == 5 ? 1 : (q == 6 ? -1 : 0); }
template <typename R>
__host__ __device__ constexpr R TW(int q) {
  return q == 0 ? R(0.25) : R(0.125);
}

template <typename R>
__device__ inline R omega_of(const ThermalArgs& a) {
  if constexpr (sizeof(R) == 4) return a.omega_f;
  else return a.omega_d;
}

__device__ inline void decompose(const ThermalArgs& a, uint32_t i, uint32_t& x, uint32_t& y, uint32_t& z) {
  const uint32_t yz = fdiv(i, a.dnx);
  x = i - yz * a.nx;
  z = fdiv(yz, a.dny);
  y = yz - z * a.ny;
}

// WT: also store T = sum_q g_q of the input state into its temperature slot
// (a.T) -- the value the previous step's temperature pass would have
// written, so a run of pure-diffusion steps needs that pass only once.
template <typename R, bool CONV, bool WT = false>
__global__ void __launch_bounds__(256) k_thermal_collide(const ThermalArgs a) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= a.N) return;
  uint32_t x, y, z;
  decompose(a, i, x, y, z);
  const R* __restrict__ gin = (const R*)a.gin;
  R* __restrict__ gout = (R*)a.gout;
  const size_t N = a.N;
  R g[TQ];
#pragma unroll
  for (int q = 0; q < TQ; ++q) g[q] = gin[q * N + i];
  R T = R(0);
#pragma unroll
  for (int q = 0; q < TQ; ++q) T = T + g[q];
  if constexpr (WT) ((R*)a.T)[i] = T;
  R* __restrict__ hs = (R*)a.hs;
  const R* __restrict__ hsb = (const R*)a.hsb;
  R S = hsb ? hsb[i] : hs[i];
  bool write_s = hsb != nullptr;
  if constexpr (CONV) {
    if (x >= 1 && x + 1 < a.nx && y >= 1 && y + 1 < a.ny && z >= 1 && z + 1 < a.nz) {
      const R* __restrict__ Tf = (const R*)a.T;
      const R* __restrict__ u = (const R*)a.u;
      const R dTdx = (Tf[i + 1] - Tf[i - 1]) / R(2);
      const R dTdy = (Tf[i + a.nx] - Tf[i - a.nx]) / R(2);
      const R dTdz = (Tf[i + a.nxy] - Tf[i - a.nxy]) / R(2);
      const R conv = -((u[i] * dTdx + u[N + i] * dTdy) + u[2 * N + i] * dTdz);
      S = S + conv;
      write_s = true;
    }
  }
  if (write_s) hs[i] = S;
  const R om = omega_of<R>(a);
#pragma unroll
  for (int q = 0; q < TQ; ++q) {
    const R w = TW<R>(q);
    const R geq = w * T;
    const R gs = (g[q] - om * (g[q] - geq)) + (w * S) * R(1);
    const int ex = TEX(q), ey = TEY(q), ez = TEZ(q);
    // destination x + e_q when inside the box
    const bool to_in = (ex == 0 || (ex > 0 ? x + 1 < a.nx : x >= 1)) && (ey == 0 || (ey > 0 ? y + 1 < a.ny : y >= 1)) &&
                       (ez == 0 || (ez > 0 ? z + 1 < a.nz : z >= 1));
    // own slot when the source x - e_q lies outside the box
    const bool from_out = (ex > 0 && x == 0) || (ex < 0 && x + 1 == a.nx) || (ey > 0 && y == 0) ||
                          (ey < 0 && y + 1 == a.ny) || (ez > 0 && z == 0) || (ez < 0 && z + 1 == a.nz);
    const int64_t off = (int64_t)ex + (int64_t)ey * a.nx + (int64_t)ez * a.nxy;
    if (q == 0) {
      gout[i] = gs;
    } else {
      if (to_in) gout[q * N + (size_t)((int64_t)i + off)] = gs;
      if (from_out) gout[q * N + i] = gs;
    }
  }
}

// T' = sum_q g'_q and the stability test, per-block counts of failing cells.
template <typename R>
__global__ void __launch_bounds__(256) k_thermal_temperature(const ThermalArgs a) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  double bad = 0.0;
  if (i < a.N) {
    const R* __restrict__ g = (const R*)a.gout;
    const size_t N = a.N;
    R T = R(0);
    bool gbad = false;
#pragma unroll
    for (int q = 0; q < TQ; ++q) {
      const R v = g[q * N + i];
      T = T + v;
      gbad = gbad || !(v == v) || fabs(v) > R(1000);
    }
    ((R*)a.Tout)[i] = T;
    if (T < R(-50) || T > R(150) || !(T == T) || gbad) bad = 1.0;
  }
  bad = wave_sum(bad);
  __shared__ double sh[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = bad;
  __syncthreads();
  if (threadIdx.x == 0) a.partial[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ void __launch_bounds__(256) k_thermal_fold(const double* __restrict__ partial, uint32_t nblk, int nval,
                                                      double* __restrict__ out) {
  // nval = 1: sum; nval = 3: {min, max, sum} triples
  __shared__ double sh[4][3];
  double v[3] = {1.0e300, -1.0e300, 0.0};
  if (nval == 1) v[2] = 0.0;
  for (uint32_t b = threadIdx.x; b < nblk; b += 256) {
    if (nval == 1) {
      v[2] += partial[b];
    } else {
      v[0] = fmin(v[0], partial[(size_t)b * 3]);
      v[1] = fmax(v[1], partial[(size_t)b * 3 + 1]);
      v[2] += partial[(size_t)b * 3 + 2];
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  v[0] = wave_min(v[0]);
  v[1] = wave_max(v[1]);
  v[2] = wave_sum(v[2]);
  if (lane == 0) {
    sh[wid][0] = v[0];
    sh[wid][1] = v[1];
    sh[wid][2] = v[2];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r0 = sh[0][0], r1 = sh[0][1], r2 = sh[0][2];
    for (int w = 1; w < 4; ++w) {
      r0 = fmin(r0, sh[w][0]);
      r1 = fmax(r1, sh[w][1]);
      r2 += sh[w][2];
    }
    if (nval == 1) {
      out[0] = r2;
    } else {
      out[0] = r0;
      out[1] = r1;
      out[2] = r2;
    }
  }
}

template <typename R>
__global__ void __launch_bounds__(256) k_thermal_stats(const R* __restrict__ T, uint32_t N, double* __restrict__ partial) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  double mn = 1.0e300, mx = -1.0e300, sm = 0.0;
  if (i < N) {
    const double t = (double)T[i];
    mn = t;
    mx = t;
    sm = t;
  }
  mn = wave_min(mn);
  mx = wave_max(mx);
  sm = wave_sum(sm);
  __shared__ double sh[4][3];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    sh[wid][0] = mn;
    sh[wid][1] = mx;
    sh[wid][2] = sm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r0 = sh[0][0], r1 = sh[0][1], r2 = sh[0][2];
    for (int w = 1; w < 4; ++w) {
      r0 = fmin(r0, sh[w][0]);
      r1 = fmax(r1, sh[w][1]);
      r2 += sh[w][2];
    }
    partial[(size_t)blockIdx.x * 3] = r0;
    partial[(size_t)blockIdx.x * 3 + 1] = r1;
    partial[(size_t)blockIdx.x * 3 + 2] = r2;
  }
}

template <typename R>
__global__ void __launch_bounds__(256) k_thermal_flux(const ThermalArgs a, R* __restrict__ flux) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= a.N) return;
  uint32_t x, y, z;
  decompose(a, i, x, y, z);
  const size_t N = a.N;
  R q0 = R(0), q1 = R(0), q2 = R(0);
  if (x >= 1 && x + 1 < a.nx && y >= 1 && y + 1 < a.ny && z >= 1 && z + 1 < a.nz) {
    const R* __restrict__ Tf = (const R*)a.T;
    const R dTdx = (Tf[i + 1] - Tf[i - 1]) / R(2);
    const R dTdy = (Tf[i + a.nx] - Tf[i - a.nx]) / R(2);
    const R dTdz = (Tf[i + a.nxy] - Tf[i - a.nxy]) / R(2);
    const R k = a.k ? ((const R*)a.k)[i] : R(0.68);
    q0 = -k * dTdx;
    q1 = -k * dTdy;
    q2 = -k * dTdz;
  }
  flux[i] = q0;
  flux[N + i] = q1;
  flux[2 * N + i] = q2;
}

template <typename R>
__global__ void __launch_bounds__(256) k_thermal_init(const ThermalArgs a) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= a.N) return;
  const R T = ((const R*)a.T)[i];
  R* __restrict__ g = (R*)a.gout;
#pragma unroll
  for (int q = 0; q < TQ; ++q) g[(size_t)q * a.N + i] = TW<R>(q) * T;
}

template <typename R>
__global__ void __launch_bounds__(256) k_thermal_boundary(const ThermalArgs a, int kind, const uint8_t* __restrict__ mask,
                                                          R va, R vb) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= a.N || !mask[i]) return;
  if (kind == 1) {
    ((R*)a.Tout)[i] = va;
    R* __restrict__ g = (R*)a.gout;
#pragma unroll
    for (int q = 0; q < TQ; ++q) g[(size_t)q * a.N + i] = TW<R>(q) * va;
  } else if (kind == 2) {
    ((R*)a.hs)[i] = va;
  } else {
    const R Ts = ((const R*)a.T)[i];
    ((R*)a.hs)[i] = -(va * (Ts - vb));
  }
}

// ---- launchers ----------------------------------------------------------------
template <typename R>
static hipError_t thermal_steps(const ThermalArgs& base, void* const g[2], void* const T[2], int& gcur, int& tcur,
                                int nsteps, hipStream_t s) {
  // With convection a step's source needs the whole temperature field of its
  // input state before it starts, so every step runs the temperature pass.
  // Pure diffusion reads only the cell's own populations: the collide pass of
  // step n + 1 stores T of its input state (state n + 1) itself, and the
  // temperature pass (with the stability test that is reported) runs after
  // the last step only.  The first step leaves the input state's temperature
  // as the caller set it (it may be a boundary value rather than sum g).
  const dim3 grid((base.N + 255) / 256);
  for (int n = 0; n < nsteps; ++n) {
    ThermalArgs a = base;
    a.gin = g[gcur];
    a.gout = g[1 - gcur];
    a.T = T[tcur];
    a.Tout = T[1 - tcur];
    const bool last = n + 1 == nsteps;
    if (a.conv)
      hipLaunchKernelGGL((k_thermal_collide<R, true>), grid, dim3(256), 0, s, a);
    else if (n == 0)
      hipLaunchKernelGGL((k_thermal_collide<R, false>), grid, dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((k_thermal_collide<R, false, true>), grid, dim3(256), 0, s, a);
    if (a.conv || last) hipLaunchKernelGGL(k_thermal_temperature<R>, grid, dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    gcur = 1 - gcur;
    tcur = 1 - tcur;
  }
  return hipSuccess;
}

hipError_t launch_thermal_steps(bool dbl, const ThermalArgs& base, void* const g[2], void* const T[2], int& gcur,
                                int& tcur, int nsteps, hipStream_t s) {
  return dbl ? thermal_steps<double>(base, g, T, gcur, tcur, nsteps, s)
             : thermal_steps<float>(base, g, T, gcur, tcur, nsteps, s);
}

hipError_t launch_thermal_fold(const double* partial, uint32_t nblk, int nval, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_thermal_fold, dim3(1), dim3(256), 0, s, partial, nblk, nval, out);
  return hipGetLastError();
}

hipError_t launch_thermal_stats(bool dbl, const void* T, uint32_t N, double* partial, hipStream_t s) {
  const dim3 grid((N + 255) / 256);
  if (dbl)
    hipLaunchKernelGGL(k_thermal_stats<double>, grid, dim3(256), 0, s, (const double*)T, N, partial);
  else
    hipLaunchKernelGGL(k_thermal_stats<float>, grid, dim3(256), 0, s, (const float*)T, N, partial);
  return hipGetLastError();
}

hipError_t launch_thermal_flux(bool dbl, const ThermalArgs& a, void* flux, hipStream_t s) {
  const dim3 grid((a.N + 255) / 256);
  if (dbl)
    hipLaunchKernelGGL(k_thermal_flux<double>, grid, dim3(256), 0, s, a, (double*)flux);
  else
    hipLaunchKernelGGL(k_thermal_flux<float>, grid, dim3(256), 0, s, a, (float*)flux);
  return hipGetLastError();
}

hipError_t launch_thermal_init(bool dbl, const ThermalArgs& a, hipStream_t s) {
  const dim3 grid((a.N + 255) / 256);
  if (dbl)
    hipLaunchKernelGGL(k_thermal_init<double>, grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_thermal_init<float>, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_thermal_boundary(bool dbl, const ThermalArgs& a, int kind, const uint8_t* mask, double va, double vb,
                                   hipStream_t s) {
  const dim3 grid((a.N + 255) / 256);
  if (dbl)
    hipLaunchKernelGGL(k_thermal_boundary<double>, grid, dim3(256), 0, s, a, kind, mask, va, vb);
  else
    hipLaunchKernelGGL(k_thermal_boundary<float>, grid, dim3(256), 0, s, a, kind, mask, (float)va, (float)vb);
  return hipGetLastError();
}

}  // namespace pol
