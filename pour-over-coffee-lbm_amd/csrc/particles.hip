// Lagrangian coffee-particle coupling kernels (one thread per particle),
// templated on the real type (float for F32/F16/BF16 fluid storage, double
// for the FP64 oracle run).
//
// couple : particle -> cell index + clamped trilinear fractions, gather of the
//          fluid velocity at the 8 corners, Schiller-Naumann drag
//          drag = beta (u_f - v), drag under-relaxation, and the reaction
//          spread to the 8 corners with hardware floating-point atomics as
//          the pair (w beta v, w beta) into the particle-drag field; the next
//          stream-collide step applies sum w beta (v - u) with the cell
//          velocity u implicit (the default).  The reference's explicit form
//          -- the reaction -drag spread into a [3][N] reaction field that the
//          caller adds to the body force, coffee_particles.py:1146-1149 and
//          legacy/lbm_solver.py:1478-1483 -- is the second mode (ParticleArgs
//          reaction set).  With the lognormal size tail of the brew workload,
//          particles up to 6 cells wide give beta ~ 3 per cell and the
//          explicit reaction diverges within ~30 steps, hence the implicit
//          default.
// advance: zero this step's 8-corner footprint in the drag field (so the
//          field never needs a full-grid clear), then move the particle
//          (linearised-implicit drag, buoyancy-corrected gravity, displacement
//          clamp, rejection at walls and at the filter layer, removal at
//          the outlet).
//
// Reference (coffee_particles.py): index/weights :1159-1174 (dupes
// :1008-1022, :1052-1066), gather :1176-1196, drag coefficient :1088-1099,
// drag + scatter :1107-1154 and :1048-1086, under-relaxation :1200-1212,
// clear :1101-1105, particle update :641-720.
#include "engine.h"

namespace pol {

// i = (int)max(0, min(n - 2, p)) with min/max in the position's type (the
// reference's ti.min/ti.max promote the int bound to f32) and truncation,
// frac = clamp(p - i, 0, 1).  fmin/fmax return the non-NaN operand, so a NaN
// coordinate maps to n-2 with fraction 1; the trailing + 0 turns a -0 fraction
// (p = -0) into +0 whatever sign max(0, -0) returns.
__device__ __forceinline__ void cell_index(float p, int n, int& i, float& fr) {
  const float c = fmaxf(0.0f, fminf((float)(n - 2), p));
  i = (int)c;
  const float t = p - (float)i;
  fr = fmaxf(0.0f, fminf(1.0f, t)) + 0.0f;
}
__device__ __forceinline__ void cell_index(double p, int n, int& i, double& fr) {
  const double c = fmax(0.0, fmin((double)(n - 2), p));
  i = (int)c;
  const double t = p - (double)i;
  fr = fmax(0.0, fmin(1.0, t)) + 0.0;
}

// Schiller-Naumann with the reference's piecewise bounds:
//   Re < 0.1        : 24 / max(0.01, Re)
//   0.1 <= Re < 1000: 24/Re (1 + 0.15 Re^0.687)
//   Re >= 1000      : 0.44
template <typename R>
__device__ __forceinline__ R drag_coefficient(R re) {
  R cd = R(24) / fmax(R(0.01), re);
  if (re >= R(0.1) && re < R(1000)) cd = R(24) / re * (R(1) + R(0.15) * pow(re, R(0.687)));
  else if (re >= R(1000)) cd = R(0.44);
  return cd;
}

__device__ __forceinline__ void atomic_add(float* p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }

// Eight lanes per particle, one per trilinear corner c = 4 dx + 2 dy + dz:
// a particle's gather, scatter and footprint clear run as eight independent
// memory streams, so a 50 000-particle launch keeps 6 250 waves in flight
// instead of 782 (the one-thread-per-particle form was latency bound at
// ~100 us per coupling).  The eight partial velocities meet in a butterfly
// over lanes c ^ 1, c ^ 2, c ^ 4, which leaves the same sum in all eight.
constexpr int CORNERS = 8;

template <typename R>
__device__ __forceinline__ R corner_weight(int c, R fx, R fy, R fz) {
  const R wx = (c & 4) ? fx : R(1) - fx, wy = (c & 2) ? fy : R(1) - fy, wz = (c & 1) ? fz : R(1) - fz;
  return wx * wy * wz;
}

// EXPLICIT: scatter the reaction -drag_new (coffee_particles.py:1146-1149,
// 1049-1086) into the [3][N] reaction field instead of the implicit drag
// field.  RELAX: under-relax in the same pass (the native coupled loop);
// the reference applies it as a separate call (apply_under_relaxation).
template <typename R, bool EXPLICIT, bool RELAX>
__global__ void __launch_bounds__(256) k_particles_couple(const ParticleArgs a) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int p = t >> 3, c = t & 7;
  if (p >= a.count) return;  // whole 8-lane groups
  const int32_t* active = a.active;
  if (!active[p]) return;    // whole 8-lane groups
  const R* pos = (const R*)a.pos;
  const R* vel = (const R*)a.vel;
  const R* ufield = (const R*)a.u_field;
  R* field = (R*)(EXPLICIT ? a.reaction : a.drag_field);
  const int P = a.count;
  const R px = pos[p], py = pos[P + p], pz = pos[2 * P + p];
  int i, j, k;
  R fx, fy, fz;
  cell_index(px, (int)a.nx, i, fx);
  cell_index(py, (int)a.ny, j, fy);
  cell_index(pz, (int)a.nz, k, fz);
  const R w = corner_weight<R>(c, fx, fy, fz);
  const uint32_t nx = a.nx, nxy = a.nxy, N = a.N;
  const uint32_t cell = ((uint32_t)k + (c & 1)) * nxy + ((uint32_t)j + ((c >> 1) & 1)) * nx + (uint32_t)i + (c >> 2);
  R uf[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    R v = w * ufield[(size_t)d * N + cell];
    v += __shfl_xor(v, 1, CORNERS);
    v += __shfl_xor(v, 2, CORNERS);
    v += __shfl_xor(v, 4, CORNERS);
    uf[d] = v;
  }
  const R rx = uf[0] - vel[p], ry = uf[1] - vel[P + p], rz = uf[2] - vel[2 * P + p];
  const R mag = sqrt(rx * rx + ry * ry + rz * rz);
  R dn[3] = {R(0), R(0), R(0)};
  R re = R(0), cd = R(0);
  if (mag > R(1e-8)) {
    const R r = ((const R*)a.radius)[p];
    const R rho_f = (R)a.rho_fluid;
    const R mu = rho_f * (R)a.nu_fluid;
    re = rho_f * mag * R(2) * r / fmax(R(1e-8), mu);
    cd = drag_coefficient<R>(re);
    const R area = R(3.14159) * r * r;  // the reference's literal
    R dmag = R(0.5) * rho_f * cd * area * mag;
    if (a.drag_law == POL_DRAG_TEXTBOOK) dmag *= mag;
    dmag = fmin(dmag, ((const R*)a.mass)[p] * (R)a.drag_max_per_mass);
    const R beta = dmag / mag;  // drag = beta (u_f - v)
    dn[0] = beta * rx;
    dn[1] = beta * ry;
    dn[2] = beta * rz;
    if constexpr (EXPLICIT) {
      // Newton's third law as the reference writes it: corner c receives
      // w_c (-drag_new)
#pragma unroll
      for (int d = 0; d < 3; ++d) atomic_add(field + (size_t)d * N + cell, w * (-dn[d]));
    } else {
      // Newton's third law, fluid side implicit: the corner receives w beta v
      // and w beta; the collision applies sum w beta (v - u_cell) with u_cell
      // solved implicitly (cell_model.h)
      const R comp[4] = {beta * vel[p], beta * vel[P + p], beta * vel[2 * P + p], beta};
#pragma unroll
      for (int d = 0; d < 4; ++d) atomic_add(field + (size_t)d * N + cell, w * comp[d]);
    }
  }
  if (c != 0) return;
  ((R*)a.re_p)[p] = re;
  ((R*)a.cd)[p] = cd;
  const R al = (R)a.alpha;
  R* drag_new = (R*)a.drag_new;
  R* drag_old = (R*)a.drag_old;
  R* drag = (R*)a.drag;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    ((R*)a.u_fluid)[(size_t)d * P + p] = uf[d];
    drag_new[(size_t)d * P + p] = dn[d];
    if constexpr (RELAX) {
      const R rel = al * dn[d] + (R(1) - al) * drag_old[(size_t)d * P + p];
      drag[(size_t)d * P + p] = rel;
      drag_old[(size_t)d * P + p] = rel;
    }
  }
}

// apply_under_relaxation (coffee_particles.py:1200-1212):
// drag = alpha drag_new + (1 - alpha) drag_old, drag_old = drag, active only.
template <typename R>
__global__ void __launch_bounds__(256) k_particles_relax(const ParticleArgs a, R al) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= a.count || !a.active[p]) return;
  const int P = a.count;
  const R* drag_new = (const R*)a.drag_new;
  R* drag_old = (R*)a.drag_old;
  R* drag = (R*)a.drag;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const R rel = al * drag_new[(size_t)d * P + p] + (R(1) - al) * drag_old[(size_t)d * P + p];
    drag[(size_t)d * P + p] = rel;
    drag_old[(size_t)d * P + p] = rel;
  }
}

// Zero the 8-corner drag-field footprint of each active particle at its
// current position (one lane per corner), the cells its last coupling pass
// wrote.
template <typename R>
__global__ void __launch_bounds__(256) k_particles_clear_footprint(const ParticleArgs a) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int p = t >> 3, corner = t & 7;
  if (p >= a.count || !a.active[p]) return;
  const R* pos = (const R*)a.pos;
  const int P = a.count;
  int i, j, k;
  R tt;
  cell_index(pos[p], (int)a.nx, i, tt);
  cell_index(pos[P + p], (int)a.ny, j, tt);
  cell_index(pos[2 * P + p], (int)a.nz, k, tt);
  const uint32_t cell = ((uint32_t)k + (corner & 1)) * a.nxy + ((uint32_t)j + ((corner >> 1) & 1)) * a.nx +
                        (uint32_t)i + (corner >> 2);
  R* field = (R*)a.drag_field;
#pragma unroll
  for (int d = 0; d < 4; ++d) field[(size_t)d * a.N + cell] = R(0);
}

template <typename R>
__global__ void __launch_bounds__(256) k_zero_real(R* __restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = R(0);
}

// add_particle_reaction_forces (legacy/lbm_solver.py:1478-1483): the body
// force takes the reaction on every non-solid cell.
template <typename R>
__global__ void __launch_bounds__(256) k_add_field_fluid(R* __restrict__ force, const R* __restrict__ add,
                                                         const uint32_t* __restrict__ flags, uint32_t N) {
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  if (c >= N) return;
  if (flags && ((flags[c] >> POL_KIND_SHIFT) & POL_KIND_MASK) == POL_KIND_SOLID) return;
#pragma unroll
  for (int d = 0; d < 3; ++d) force[(size_t)d * N + c] += add[(size_t)d * N + c];
}

template <typename R>
__global__ void __launch_bounds__(256) k_particles_advance(const ParticleArgs a) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int p = t >> 3, corner = t & 7;
  if (p >= a.count) return;
  int32_t* active = a.active;
  if (!active[p]) return;
  R* pos = (R*)a.pos;
  R* vel = (R*)a.vel;
  R* field = (R*)a.drag_field;
  const int P = a.count;
  const uint32_t nx = a.nx, ny = a.ny, nz = a.nz, nxy = a.nxy, N = a.N;
  const R x[3] = {pos[p], pos[P + p], pos[2 * P + p]};
  // 1. clear the drag-field footprint written by couple (same corner, one per lane)
  {
    int i, j, k;
    R tt;
    cell_index(x[0], (int)nx, i, tt);
    cell_index(x[1], (int)ny, j, tt);
    cell_index(x[2], (int)nz, k, tt);
    const uint32_t cell = ((uint32_t)k + (corner & 1)) * nxy + ((uint32_t)j + ((corner >> 1) & 1)) * nx +
                          (uint32_t)i + (corner >> 2);
    if (field) {  // implicit drag field only (uniform)
#pragma unroll
      for (int d = 0; d < 4; ++d) field[(size_t)d * N + cell] = R(0);
    }
  }
  if (corner != 0) return;
  // 2. velocity: drag linearised about the current relative velocity,
  //    dv/dt = beta/m (u_f - v) + g, implicit in beta (stable for any
  //    particle response time)
  const R* ufl = (const R*)a.u_fluid;
  const R* dnp = (const R*)a.drag_new;
  R v[3] = {vel[p], vel[P + p], vel[2 * P + p]};
  const R uf[3] = {ufl[p], ufl[P + p], ufl[2 * P + p]};
  const R dn[3] = {dnp[p], dnp[P + p], dnp[2 * P + p]};
  const R rx = uf[0] - v[0], ry = uf[1] - v[1], rz = uf[2] - v[2];
  const R mag = sqrt(rx * rx + ry * ry + rz * rz);
  const R dmag = sqrt(dn[0] * dn[0] + dn[1] * dn[1] + dn[2] * dn[2]);
  const R m = ((const R*)a.mass)[p];
  const R beta = mag > R(1e-8) ? dmag / mag : R(0);
  const R dt = (R)a.dt;
  const R h = dt * beta / m;
#pragma unroll
  for (int d = 0; d < 3; ++d) v[d] = (v[d] + h * uf[d] + dt * (R)a.g[d]) / (R(1) + h);
  // 3. position with displacement clamp
  R dx[3] = {dt * v[0], dt * v[1], dt * v[2]};
  const R dl = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
  const R max_disp = (R)a.max_disp;
  if (dl > max_disp) {
    const R s = max_disp / dl;
    dx[0] *= s;
    dx[1] *= s;
    dx[2] *= s;
  }
  R xn[3] = {x[0] + dx[0], x[1] + dx[1], x[2] + dx[2]};
  // 4. boundaries: leave through the outlet / bottom; reject moves into walls
  //    or the filter layer (the paper retains the grounds)
  if (!(xn[2] >= R(1)) || !(xn[0] >= R(0)) || !(xn[1] >= R(0)) || !(xn[0] < (R)(nx - 1)) ||
      !(xn[1] < (R)(ny - 1)) || !(xn[2] < (R)(nz - 1))) {
    active[p] = 0;
  } else {
    const uint32_t c = (uint32_t)xn[2] * nxy + (uint32_t)xn[1] * nx + (uint32_t)xn[0];
    const uint32_t fl = a.flags ? a.flags[c] : 0u;
    const uint32_t kd = (fl >> POL_KIND_SHIFT) & POL_KIND_MASK;
    if (kd == POL_KIND_OUTLET) {
      active[p] = 0;
    } else if (kd == POL_KIND_SOLID || (fl & POL_FLAG_POROUS)) {
      const R wd = (R)a.wall_damping;
      for (int d = 0; d < 3; ++d) {
        xn[d] = x[d];
        v[d] *= wd;
      }
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    pos[(size_t)d * P + p] = xn[d];
    vel[(size_t)d * P + p] = v[d];
  }
}

__global__ void __launch_bounds__(256) k_particle_index(const float* __restrict__ pos, int32_t n, int32_t nx,
                                                        int32_t ny, int32_t nz, int32_t* __restrict__ idx,
                                                        float* __restrict__ frac) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const int dims[3] = {nx, ny, nz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int i;
    float t;
    cell_index(pos[(size_t)d * n + p], dims[d], i, t);
    idx[(size_t)d * n + p] = i;
    frac[(size_t)d * n + p] = t;
  }
}

template <typename R>
static void launch_couple_t(const ParticleArgs& a, bool relax, hipStream_t s) {
  const dim3 grid(((size_t)a.count * CORNERS + 255) / 256), block(256);
  if (a.reaction) {
    if (relax) hipLaunchKernelGGL((k_particles_couple<R, true, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_particles_couple<R, true, false>), grid, block, 0, s, a);
  } else {
    if (relax) hipLaunchKernelGGL((k_particles_couple<R, false, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_particles_couple<R, false, false>), grid, block, 0, s, a);
  }
}
hipError_t launch_particles_couple(const ParticleArgs& a, bool relax, hipStream_t s) {
  if (a.count <= 0) return hipSuccess;
  if (a.real_is_double) launch_couple_t<double>(a, relax, s);
  else launch_couple_t<float>(a, relax, s);
  return hipGetLastError();
}
hipError_t launch_particles_relax(const ParticleArgs& a, double alpha, hipStream_t s) {
  if (a.count <= 0) return hipSuccess;
  const dim3 grid(((size_t)a.count + 255) / 256), block(256);
  if (a.real_is_double) hipLaunchKernelGGL(k_particles_relax<double>, grid, block, 0, s, a, alpha);
  else hipLaunchKernelGGL(k_particles_relax<float>, grid, block, 0, s, a, (float)alpha);
  return hipGetLastError();
}
hipError_t launch_particles_clear_footprint(const ParticleArgs& a, hipStream_t s) {
  if (a.count <= 0 || !a.drag_field) return hipSuccess;
  const dim3 grid(((size_t)a.count * CORNERS + 255) / 256), block(256);
  if (a.real_is_double) hipLaunchKernelGGL(k_particles_clear_footprint<double>, grid, block, 0, s, a);
  else hipLaunchKernelGGL(k_particles_clear_footprint<float>, grid, block, 0, s, a);
  return hipGetLastError();
}
hipError_t launch_zero_real(bool dbl, void* p, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const dim3 grid((n + 255) / 256), block(256);
  if (dbl) hipLaunchKernelGGL(k_zero_real<double>, grid, block, 0, s, (double*)p, n);
  else hipLaunchKernelGGL(k_zero_real<float>, grid, block, 0, s, (float*)p, n);
  return hipGetLastError();
}
hipError_t launch_add_field_fluid(bool dbl, void* force, const void* add, const uint32_t* flags, uint32_t N,
                                  hipStream_t s) {
  if (N == 0) return hipSuccess;
  const dim3 grid((N + 255) / 256), block(256);
  if (dbl) hipLaunchKernelGGL(k_add_field_fluid<double>, grid, block, 0, s, (double*)force, (const double*)add, flags, N);
  else hipLaunchKernelGGL(k_add_field_fluid<float>, grid, block, 0, s, (float*)force, (const float*)add, flags, N);
  return hipGetLastError();
}
hipError_t launch_particles_advance(const ParticleArgs& a, hipStream_t s) {
  if (a.count <= 0) return hipSuccess;
  if (a.real_is_double)
    hipLaunchKernelGGL(k_particles_advance<double>, dim3(((size_t)a.count * CORNERS + 255) / 256), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_particles_advance<float>, dim3(((size_t)a.count * CORNERS + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_particle_index(const float* pos, int32_t n, int32_t nx, int32_t ny, int32_t nz, int32_t* idx,
                                 float* frac, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_particle_index, dim3((n + 255) / 256), dim3(256), 0, s, pos, n, nx, ny, nz, idx, frac);
  return hipGetLastError();
}

}  // namespace pol
