// Phase-field (Cahn-Hilliard + CSF surface tension) producer for gfx950.
//
// Replaces (reference): src/core/multiphase_3d.py as it resolves at run time
//   :78-108  compute_chemical_potential (laplacian_phi, mu)
//   :110-131 compute_gradients (grad_phi, grad_mu, normal)
//   :133-149 compute_curvature (div n)
//   :365-381 compute_surface_tension_force (the later definition wins:
//            sigma kappa |grad phi| n where |phi| < 0.9)
//   :403-411 apply_surface_tension (body_force += F / rho, fluid cells)
//   :151-197 update_phase_field_cahn_hilliard, :383-401 apply_phase_separation,
//   :418-422 copy_phase_field, :413-428 update_density_from_phase
//   :389-414 step, :416-421 accumulate_surface_tension_pre_collision
//
// The reference runs up to eight grid passes through intermediate fields
// (grad_phi, normal, curvature, surface_force, phi_new).  Here:
//   k_mp_apply: the pre-collision pass.  Gradients, normal, curvature and the
//     CSF force of a cell are formed in registers -- the neighbours' normals
//     from the 25-point phi stencil instead of a stored normal field -- and
//     only where a force can arise (fluid interior cells with |phi| < 0.9);
//     then body_force += F / rho.
//   k_mp_interface: the same per-cell forms for every interior cell, written
//     out as the reference's grad_phi / grad_mu / normal / curvature /
//     surface_force fields when the caller asks for them.
//   k_mp_phase: upwind advection, M lap(mu), clamp, phase separation, the
//     copy to phi (ping-pong) and the density / phase update in one pass.
// Face cells are never written by the reference's step: their normal and
// surface force stay 0 and the step copies their phi_new values (held in
// the phi_new buffer) into phi.
//
// Arithmetic: float32 in the reference's operation order with contraction
// off, correctly rounded division and square root, so a float32 numpy
// restatement of the same expressions agrees bit for bit.
#include "engine.h"
#include "phase_cell.h"

namespace pol {

template <typename R>
__device__ inline float ld_real(const void* p, size_t i) {
  return (float)((const R*)p)[i];
}

// Pre-collision pass without stored fields (apply_surface_tension of the
// current phi).  Only fluid interior cells with |phi| < 0.9 can carry a
// force; every other fluid cell adds +0 / rho, which changes its body force
// only where a component is -0.0 (it becomes +0.0 when rho > 1e-10), so those
// cells read the body force and write nothing else.
template <typename R>
__global__ void __launch_bounds__(256) k_mp_apply(const MultiphaseArgs a) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= a.N) return;
  if (a.flags && ((a.flags[i] >> POL_KIND_SHIFT) & POL_KIND_MASK) == POL_KIND_SOLID) return;
  uint32_t x, y, z;
  mp_decompose(a, i, x, y, z);
  const size_t N = a.N;
  R* __restrict__ bf = (R*)a.force;
  const R b0 = bf[i], b1 = bf[N + i], b2 = bf[2 * N + i];
  float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
  bool force = false;
  if (mp_interior(a, x, y, z)) {
    const float p = a.phi[i];
    if (fabsf(p) < 0.9f) {
      const CellForce c = cell_force(a, i, x, y, z, p);
      f0 = c.f0;
      f1 = c.f1;
      f2 = c.f2;
      force = true;
    }
  }
  const bool negzero = (b0 == R(0) && signbit(b0)) || (b1 == R(0) && signbit(b1)) || (b2 == R(0) && signbit(b2));
  if (!force && !negzero) return;  // b + (+0 / rho) == b bit for bit
  const R rho = ((const R*)a.rho)[i];
  if (!(rho > R(1e-10f))) return;
  bf[i] = b0 + (R)f0 / rho;
  bf[N + i] = b1 + (R)f1 / rho;
  bf[2 * N + i] = b2 + (R)f2 / rho;
}

template <typename R, bool OUT>
__global__ void __launch_bounds__(256) k_mp_interface(const MultiphaseArgs a) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= a.N) return;
  uint32_t x, y, z;
  mp_decompose(a, i, x, y, z);
  const size_t N = a.N;
  float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
  if (mp_interior(a, x, y, z)) {
    const CellForce c = cell_force(a, i, x, y, z, a.phi[i]);
    const float gx = c.gx, gy = c.gy, gz = c.gz, n0 = c.n0, n1 = c.n1, n2 = c.n2, kap = c.kap;
    f0 = c.f0;
    f1 = c.f1;
    f2 = c.f2;
    if constexpr (OUT) {
      if (a.grad_phi) {
        a.grad_phi[i] = gx;
        a.grad_phi[N + i] = gy;
        a.grad_phi[2 * N + i] = gz;
      }
      if (a.grad_mu) {
        const float* __restrict__ mu = a.mu;
        a.grad_mu[i] = (mu[i + 1] - mu[i - 1]) * 0.5f;
        a.grad_mu[N + i] = (mu[i + a.nx] - mu[i - a.nx]) * 0.5f;
        a.grad_mu[2 * N + i] = (mu[i + a.nxy] - mu[i - a.nxy]) * 0.5f;
      }
      if (a.normal) {
        a.normal[i] = n0;
        a.normal[N + i] = n1;
        a.normal[2 * N + i] = n2;
      }
      if (a.curvature) a.curvature[i] = kap;
      if (a.sforce) {
        a.sforce[i] = f0;
        a.sforce[N + i] = f1;
        a.sforce[2 * N + i] = f2;
      }
    }
  }
  if (a.apply) {
    const bool solid = a.flags && ((a.flags[i] >> POL_KIND_SHIFT) & POL_KIND_MASK) == POL_KIND_SOLID;
    if (!solid) {
      const R rho = ((const R*)a.rho)[i];
      if (rho > R(1e-10f)) {
        R* __restrict__ bf = (R*)a.force;
        bf[i] = bf[i] + (R)f0 / rho;
        bf[N + i] = bf[N + i] + (R)f1 / rho;
        bf[2 * N + i] = bf[2 * N + i] + (R)f2 / rho;
      }
    }
  }
}

template <typename R>
__global__ void __launch_bounds__(256) k_mp_phase(const MultiphaseArgs a) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= a.N) return;
  uint32_t x, y, z;
  mp_decompose(a, i, x, y, z);
  const size_t N = a.N;
  float v;
  if (mp_interior(a, x, y, z)) {
    const float* __restrict__ phi = a.phi;
    const float* __restrict__ mu = a.mu;
    const float p = phi[i];
    const float pxm = phi[i - 1], pxp = phi[i + 1];
    const float pym = phi[i - a.nx], pyp = phi[i + a.nx];
    const float pzm = phi[i - a.nxy], pzp = phi[i + a.nxy];
    const float ux = ld_real<R>(a.u, i), uy = ld_real<R>(a.u, N + i), uz = ld_real<R>(a.u, 2 * N + i);
    const float dx = ux > 0.0f ? p - pxm : pxp - p;
    const float dy = uy > 0.0f ? p - pym : pyp - p;
    const float dz = uz > 0.0f ? p - pzm : pzp - p;
    const float conv = -((ux * dx + uy * dy) + uz * dz);
    const float lmu = (((((mu[i + 1] + mu[i - 1]) + mu[i + a.nx]) + mu[i - a.nx]) + mu[i + a.nxy]) + mu[i - a.nxy]) -
                      6.0f * mu[i];
    const float diff = a.mobility * lmu;
    v = p + a.dt * (conv + diff);
    v = fmaxf(-1.0f, fminf(1.0f, v));
    if (fabsf(p) < 0.99f) {
      const float lap = (((((pxp + pxm) + pyp) + pym) + pzp) + pzm) - 6.0f * p;
      const float chem = p * (p * p - 1.0f) - 0.01f * lap;
      v = v + (-0.001f * chem) * a.dt;
    }
  } else {
    v = a.phi_new[i];
  }
  a.phi_out[i] = v;
  const float pc = fmaxf(-1.0f, fminf(1.0f, v));
  ((R*)a.rho)[i] = (R)(a.rho_air + (a.drho * (pc + 1.0f)) / 2.0f);
  ((R*)a.phase)[i] = (R)((pc + 1.0f) / 2.0f);
}

// update_density_from_phase alone (rho and phase from the current phi)
template <typename R>
__global__ void __launch_bounds__(256) k_mp_density(const MultiphaseArgs a) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= a.N) return;
  const float pc = fmaxf(-1.0f, fminf(1.0f, a.phi[i]));
  ((R*)a.rho)[i] = (R)(a.rho_air + (a.drho * (pc + 1.0f)) / 2.0f);
  ((R*)a.phase)[i] = (R)((pc + 1.0f) / 2.0f);
}

__global__ void __launch_bounds__(256) k_mp_chemical_potential(const MultiphaseArgs a, float* __restrict__ lap_out) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= a.N) return;
  uint32_t x, y, z;
  mp_decompose(a, i, x, y, z);
  if (!mp_interior(a, x, y, z)) return;
  const float* __restrict__ phi = a.phi;
  const float p = phi[i];
  const float lap = (((((phi[i + 1] + phi[i - 1]) + phi[i + a.nx]) + phi[i - a.nx]) + phi[i + a.nxy]) + phi[i - a.nxy]) -
                    6.0f * p;
  if (lap_out) lap_out[i] = lap;
  a.mu[i] = (p * p * p - p) + (-a.kappa) * lap;
}

// ---- launchers ----------------------------------------------------------------
hipError_t launch_mp_interface(bool dbl, const MultiphaseArgs& a, bool out, hipStream_t s) {
  const dim3 grid((a.N + 255) / 256);
  if (!out) {  // apply only
    if (dbl) hipLaunchKernelGGL(k_mp_apply<double>, grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_mp_apply<float>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
  }
  if (dbl) {
    if (out) hipLaunchKernelGGL((k_mp_interface<double, true>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_mp_interface<double, false>), grid, dim3(256), 0, s, a);
  } else {
    if (out) hipLaunchKernelGGL((k_mp_interface<float, true>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_mp_interface<float, false>), grid, dim3(256), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_mp_phase(bool dbl, const MultiphaseArgs& a, hipStream_t s) {
  const dim3 grid((a.N + 255) / 256);
  if (dbl) hipLaunchKernelGGL(k_mp_phase<double>, grid, dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_mp_phase<float>, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_mp_density(bool dbl, const MultiphaseArgs& a, hipStream_t s) {
  const dim3 grid((a.N + 255) / 256);
  if (dbl) hipLaunchKernelGGL(k_mp_density<double>, grid, dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_mp_density<float>, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_mp_chemical_potential(const MultiphaseArgs& a, float* lap_out, hipStream_t s) {
  const dim3 grid((a.N + 255) / 256);
  hipLaunchKernelGGL(k_mp_chemical_potential, grid, dim3(256), 0, s, a, lap_out);
  return hipGetLastError();
}

}  // namespace pol
