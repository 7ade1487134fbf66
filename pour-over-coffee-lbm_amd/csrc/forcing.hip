// Body-force producers of the main loop for gfx950.
//
// Replaces (reference):
//   src/physics/precise_pouring.py:104-193  pouring region (Gaussian radial x
//     exponential vertical intensity), apply_pouring_force (downward
//     acceleration into body_force), apply_gradual_phase_change (phi -> 1)
//   src/physics/pressure_gradient_drive.py:96-279  density drive (rho pulled
//     towards the target profile), compute_pressure_gradient (-grad(rho cs2)
//     / rho, norm-limited) and its accumulation into body_force (x1 force
//     drive, x0.5 mixed drive)
//   src/physics/filter_paper.py:616-746  block_particles_at_filter (bounce
//     of particles moving down onto filter cells) and
//     update_dynamic_resistance (blockage from accumulated particles)
//
// The pouring kernels run over the bounding box of the pour region only
// (at most (2r+2)^2 x 5 cells) instead of the whole grid; every cell of the
// box applies the reference's own membership test.  The pressure-force
// kernel computes the force and adds it to the body force in one pass.
// float32 in the reference's operation order with contraction off.
#include "engine.h"
#include "phase_cell.h"

namespace pol {

// float32 exp formed in double and rounded once: the correctly rounded
// value except when the double result lies within 2^-29 of a float rounding
// boundary, and the same value numpy gives for float64 exp rounded to
// float32 -- so the pour intensity and the blockage law are reproducible
// bit for bit by a host restatement.
__device__ inline float exp_rn(float x) { return (float)exp((double)x); }

__device__ inline bool fc_solid(const ForcingArgs& f, uint32_t c) {
  return f.flags && ((f.flags[c] >> POL_KIND_SHIFT) & POL_KIND_MASK) == POL_KIND_SOLID;
}

// total intensity of the pouring region at (i, j, k) (precise_pouring.py:104-130)
__device__ inline float pour_intensity(const PourArgs& p, int i, int j, int k) {
#pragma clang fp contract(off)
  const float dx = (float)i - p.cx;
  const float dy = (float)j - p.cy;
  const float d = __builtin_sqrtf(dx * dx + dy * dy);
  float total = 0.0f;
  if (d <= p.radius && k <= p.height && (float)k >= p.zmin) {
    const float r = d / p.radius;
    const float intensity = exp_rn(-0.5f * (r * r));
    const float decay = exp_rn((float)(-(p.height - k)) / 2.0f);
    total = intensity * decay;
  }
  return total;
}

template <typename R, bool PHASE>
__global__ void __launch_bounds__(256) k_pour(const ForcingArgs f, const PourArgs p, float* __restrict__ phi) {
#pragma clang fp contract(off)
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= (uint32_t)(p.bx * p.by * p.bz)) return;
  const int i = p.x0 + (int)(t % (uint32_t)p.bx);
  const int j = p.y0 + (int)((t / (uint32_t)p.bx) % (uint32_t)p.by);
  const int k = p.z0 + (int)(t / (uint32_t)(p.bx * p.by));
  const uint32_t c = ((uint32_t)k * f.ny + (uint32_t)j) * f.nx + (uint32_t)i;
  if (fc_solid(f, c)) return;
  const float total = pour_intensity(p, i, j, k);
  if (!(total > 0.0f)) return;
  if constexpr (PHASE) {
    const float cur = phi[c];
    float rate = (1.0f - cur) / 0.05f;
    rate = fmaxf(-2.0f, fminf(2.0f, rate));
    const float change = ((rate * total) * p.dt) * p.flow_rate;
    phi[c] = fmaxf(-1.0f, fminf(1.0f, cur + change));
  } else {
    float accel = 0.0f;
    if (p.dt > 1e-8f) accel = ((p.velocity * total) * p.flow_rate) / p.dt;
    accel = fminf(accel, 10.0f);
    R* __restrict__ bf = (R*)f.force;
    const size_t N = f.N;
    bf[c] = bf[c] + R(0.0f);
    bf[N + c] = bf[N + c] + R(0.0f);
    bf[2 * N + c] = bf[2 * N + c] + (R)(-accel);
  }
}

// rho <- clamp(rho + limited(rate (target - rho))) on fluid cells
// (pressure_gradient_drive.py:96-122, mixed form :206-227)
template <typename R>
__global__ void __launch_bounds__(256) k_density_drive(const ForcingArgs f, const float* __restrict__ target,
                                                       float rate, bool half, float max_adj, float lo, float hi) {
#pragma clang fp contract(off)
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  if (c >= f.N || fc_solid(f, c)) return;
  R* __restrict__ rho = (R*)f.rho;
  const float cur = (float)rho[c];
  const float diff = target[c] - cur;
  float adj = diff * rate;
  if (half) adj = adj * 0.5f;
  if (fabsf(adj) > max_adj) adj = adj > 0.0f ? max_adj : -max_adj;
  rho[c] = (R)fmaxf(lo, fminf(hi, cur + adj));
}

// pressure_force = -grad(rho) cs2 / rho (one-sided on the faces), limited to
// max_force, on fluid cells; body_force += scale pressure_force
// (pressure_gradient_drive.py:124-173, 186-190, 276-279).  Solid cells are
// never formed and hold zero (the field starts at zero and only this pass
// writes it); a solid cell that shares its wave with fluid stores that zero
// again, so that the wave's lines of the field are written whole rather than
// read, merged and written back.
// ST: the pre-collision surface tension of the bound phase field follows in
// the same pass (multiphase_3d.py:403-411 on the sum just formed): b' = b +
// scale pf, then b' + F / rho where the separate pass would add it, so the
// result is the two passes' bit for bit with one read and one write of the
// body force instead of two.
template <typename R, bool ST>
__global__ void __launch_bounds__(256) k_pressure_force(const ForcingArgs f, const MultiphaseArgs m,
                                                        float* __restrict__ pf, float max_force, float cs2, int mode) {
#pragma clang fp contract(off)
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  if (c >= f.N) return;
  const size_t N = f.N;
  const bool solid = fc_solid(f, c);
  if (__ballot(!solid) == 0) return;  // an all-solid wave touches nothing
  if (solid) {
    pf[c] = 0.0f;
    pf[N + c] = 0.0f;
    pf[2 * N + c] = 0.0f;
    return;
  }
  R* __restrict__ bf = (R*)f.force;
  // the body force is read up front, beside the density stencil
  R b0 = R(0), b1 = R(0), b2 = R(0);
  if (mode != 0) {
    b0 = bf[c];
    b1 = bf[N + c];
    b2 = bf[2 * N + c];
  }
  const uint32_t yz = fdiv(c, f.dnx);
  const uint32_t x = c - yz * f.nx;
  const uint32_t z = fdiv(yz, f.dny);
  const uint32_t y = yz - z * f.ny;
  const R* __restrict__ rho = (const R*)f.rho;
  auto rr = [&](uint32_t j) { return (float)rho[j]; };
  const float r0 = rr(c);
  float gx, gy, gz;
  if (x > 0 && x + 1 < f.nx) gx = (rr(c + 1) - rr(c - 1)) * 0.5f;
  else if (x == 0) gx = rr(c + 1) - r0;
  else gx = r0 - rr(c - 1);
  if (y > 0 && y + 1 < f.ny) gy = (rr(c + f.nx) - rr(c - f.nx)) * 0.5f;
  else if (y == 0) gy = rr(c + f.nx) - r0;
  else gy = r0 - rr(c - f.nx);
  if (z > 0 && z + 1 < f.nz) gz = (rr(c + f.nxy) - rr(c - f.nxy)) * 0.5f;
  else if (z == 0) gz = rr(c + f.nxy) - r0;
  else gz = r0 - rr(c - f.nxy);
  const float px = gx * cs2, py = gy * cs2, pz = gz * cs2;
  float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
  if (r0 > 1e-12f) {
    f0 = -px / r0;
    f1 = -py / r0;
    f2 = -pz / r0;
    const float m = __builtin_sqrtf((f0 * f0 + f1 * f1) + f2 * f2);
    if (m > max_force) {
      const float s = max_force / m;
      f0 = f0 * s;
      f1 = f1 * s;
      f2 = f2 * s;
    }
  }
  pf[c] = f0;
  pf[N + c] = f1;
  pf[2 * N + c] = f2;
  if (mode == 0) return;
  R s0, s1, s2;
  if (mode == 1) {
    s0 = b0 + (R)f0;
    s1 = b1 + (R)f1;
    s2 = b2 + (R)f2;
  } else {
    s0 = b0 + (R)(0.5f * f0);
    s1 = b1 + (R)(0.5f * f1);
    s2 = b2 + (R)(0.5f * f2);
  }
  if constexpr (ST) {
    // k_mp_apply's cell update on s (multiphase_3d.py:365-381, 403-411)
    float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f;
    bool force = false;
    if (mp_interior(m, x, y, z)) {
      const float p = m.phi[c];
      if (fabsf(p) < 0.9f) {
        const CellForce cf = cell_force(m, c, x, y, z, p);
        t0 = cf.f0;
        t1 = cf.f1;
        t2 = cf.f2;
        force = true;
      }
    }
    const bool negzero = (s0 == R(0) && signbit(s0)) || (s1 == R(0) && signbit(s1)) || (s2 == R(0) && signbit(s2));
    const R rc = rho[c];
    if ((force || negzero) && rc > R(1e-10f)) {
      s0 = s0 + (R)t0 / rc;
      s1 = s1 + (R)t1 / rc;
      s2 = s2 + (R)t2 / rc;
    }
  }
  bf[c] = s0;
  bf[N + c] = s1;
  bf[2 * N + c] = s2;
}

// filter_paper.py:699-746 on the listed paper cells
__global__ void __launch_bounds__(256) k_filter_blockage(const uint32_t* __restrict__ cells, uint32_t ncells,
                                                         float* __restrict__ acc, float* __restrict__ blockage) {
#pragma clang fp contract(off)
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= ncells) return;
  const uint32_t c = cells[t];
  const float a = acc[c];
  const float nb = 0.9f * (1.0f - exp_rn(-0.1f * a));
  blockage[c] = 0.95f * blockage[c] + 0.05f * nb;
  acc[c] = a * 0.999f;
}

// Counter-based uniform in [0, 1): a 32-bit mix of (seed, particle, draw).
__device__ inline float fc_uniform(uint32_t seed, uint32_t p, uint32_t draw) {
  uint32_t h = seed ^ (p * 0x9E3779B9u) ^ (draw * 0x85EBCA6Bu);
  h ^= h >> 16;
  h *= 0x7FEB352Du;
  h ^= h >> 15;
  h *= 0x846CA68Bu;
  h ^= h >> 16;
  return (float)(h >> 8) * (1.0f / 16777216.0f);
}

// filter_paper.py:616-697: particles moving down next to a filter cell
// bounce (v_z -> -0.3 v_z), get a horizontal kick of +-0.005, and add 0.01
// to the cell's accumulated particles.
__global__ void __launch_bounds__(256) k_filter_block(const FilterParticleArgs a) {
#pragma clang fp contract(off)
  const uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (p >= a.count || a.active[p] == 0) return;
  const uint32_t n = a.count;
  const float px = a.pos[p], py = a.pos[n + p], pz = a.pos[2 * n + p];
  const int gx = (int)(px / a.scale), gy = (int)(py / a.scale), gz = (int)(pz / a.scale);
  if (!(gx >= 0 && gx < (int)a.nx && gy >= 0 && gy < (int)a.ny && gz >= 0 && gz < (int)a.nz)) return;
  for (int oz = -2; oz < 3; ++oz) {
    const int cz = gz + oz;
    if (cz >= 0 && cz < (int)a.nz) {
      const uint32_t c = ((uint32_t)cz * a.ny + (uint32_t)gy) * a.nx + (uint32_t)gx;
      if (a.zone[c] == 1) {
        float vz = a.vel[2 * n + p];
        if (vz < 0.0f) {
          vz = -vz * 0.3f;
          const float rx = (fc_uniform(a.seed, p, 0) - 0.5f) * 0.01f;
          const float ry = (fc_uniform(a.seed, p, 1) - 0.5f) * 0.01f;
          a.vel[p] = a.vel[p] + rx;
          a.vel[n + p] = a.vel[n + p] + ry;
          a.vel[2 * n + p] = vz;
          atomicAdd(&a.acc[c], 0.01f);
        }
        break;
      }
    }
  }
}

// ---- launchers ----------------------------------------------------------------
hipError_t launch_pour(bool dbl, const ForcingArgs& f, const PourArgs& p, float* phi, hipStream_t s) {
  const uint32_t n = (uint32_t)(p.bx * p.by * p.bz);
  if (n == 0) return hipSuccess;
  const dim3 grid((n + 255) / 256);
  if (phi) {
    if (dbl) hipLaunchKernelGGL((k_pour<double, true>), grid, dim3(256), 0, s, f, p, phi);
    else hipLaunchKernelGGL((k_pour<float, true>), grid, dim3(256), 0, s, f, p, phi);
  } else {
    if (dbl) hipLaunchKernelGGL((k_pour<double, false>), grid, dim3(256), 0, s, f, p, phi);
    else hipLaunchKernelGGL((k_pour<float, false>), grid, dim3(256), 0, s, f, p, phi);
  }
  return hipGetLastError();
}

hipError_t launch_density_drive(bool dbl, const ForcingArgs& f, const float* target, float rate, bool half,
                                float max_adj, float lo, float hi, hipStream_t s) {
  const dim3 grid((f.N + 255) / 256);
  if (dbl) hipLaunchKernelGGL(k_density_drive<double>, grid, dim3(256), 0, s, f, target, rate, half, max_adj, lo, hi);
  else hipLaunchKernelGGL(k_density_drive<float>, grid, dim3(256), 0, s, f, target, rate, half, max_adj, lo, hi);
  return hipGetLastError();
}

hipError_t launch_pressure_force(bool dbl, const ForcingArgs& f, const MultiphaseArgs* st, float* pf, float max_force,
                                 float cs2, int mode, hipStream_t s) {
  const dim3 grid((f.N + 255) / 256);
  const MultiphaseArgs m = st ? *st : MultiphaseArgs{};
  if (st) {
    if (dbl) hipLaunchKernelGGL((k_pressure_force<double, true>), grid, dim3(256), 0, s, f, m, pf, max_force, cs2, mode);
    else hipLaunchKernelGGL((k_pressure_force<float, true>), grid, dim3(256), 0, s, f, m, pf, max_force, cs2, mode);
  } else {
    if (dbl) hipLaunchKernelGGL((k_pressure_force<double, false>), grid, dim3(256), 0, s, f, m, pf, max_force, cs2, mode);
    else hipLaunchKernelGGL((k_pressure_force<float, false>), grid, dim3(256), 0, s, f, m, pf, max_force, cs2, mode);
  }
  return hipGetLastError();
}

hipError_t launch_filter_blockage(const uint32_t* cells, uint32_t ncells, float* acc, float* blockage, hipStream_t s) {
  if (ncells == 0) return hipSuccess;
  hipLaunchKernelGGL(k_filter_blockage, dim3((ncells + 255) / 256), dim3(256), 0, s, cells, ncells, acc, blockage);
  return hipGetLastError();
}

hipError_t launch_filter_block(const FilterParticleArgs& a, hipStream_t s) {
  if (a.count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_filter_block, dim3((a.count + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace pol
