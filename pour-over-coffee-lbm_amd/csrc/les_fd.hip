// The reference's finite-difference Smagorinsky eddy viscosity
// (src/physics/les_turbulence.py:318-380, _compute_sgs_from_vector), one
// thread per cell of the dense grid:
//   interior cells (1 <= i < n - 1 on every axis) with LES allowed:
//     d u_a / d x_b = (u_a(x + e_b) - u_a(x - e_b)) * 0.5
//     S11 = dudx, S22 = dvdy, S33 = dwdz, S12 = 0.5 (dudy + dvdx), ...
//     |S| = sqrt(2 (S11^2 + S22^2 + S33^2 + 2 (S12^2 + S13^2 + S23^2)))
//     nu = 0 if |S| < 1e-3 or |phi| < 0.9, else min((C_s 1)^2 |S|, 0.1)
//   every other cell: nu = 0.
// float32 in the reference's operation order with contraction off (its
// Taichi kernel evaluates each expression in f32); FP64 contexts run the same
// order in double.  The mask is the cell flags' LES-off bit (the filter
// paper, filter_paper.py:199-204, set_mask).
#include "engine.h"

namespace pol {

template <typename R>
__global__ void __launch_bounds__(256) k_les_fd(const LesFdArgs a) {
#pragma clang fp contract(off)
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  if (c >= a.N) return;
  const uint32_t yz = fdiv(c, a.dnx);
  const uint32_t x = c - yz * a.nx;
  const uint32_t z = fdiv(yz, a.dny);
  const uint32_t y = yz - z * a.ny;
  R* nu = (R*)a.nu;
  const bool interior = x >= 1 && y >= 1 && z >= 1 && x + 1 < a.nx && y + 1 < a.ny && z + 1 < a.nz;
  const bool allowed = !a.flags || !(a.flags[c] & POL_FLAG_NO_LES);
  if (!interior || !allowed) {
    nu[c] = R(0);
    return;
  }
  const R* u = (const R*)a.u;
  const size_t N = a.N;
  const uint32_t sx = 1, sy = a.nx, sz = a.nxy;
  const R h = R(0.5);
  const R dudx = (u[c + sx] - u[c - sx]) * h;
  const R dudy = (u[c + sy] - u[c - sy]) * h;
  const R dudz = (u[c + sz] - u[c - sz]) * h;
  const R dvdx = (u[N + c + sx] - u[N + c - sx]) * h;
  const R dvdy = (u[N + c + sy] - u[N + c - sy]) * h;
  const R dvdz = (u[N + c + sz] - u[N + c - sz]) * h;
  const R dwdx = (u[2 * N + c + sx] - u[2 * N + c - sx]) * h;
  const R dwdy = (u[2 * N + c + sy] - u[2 * N + c - sy]) * h;
  const R dwdz = (u[2 * N + c + sz] - u[2 * N + c - sz]) * h;
  const R S11 = dudx, S22 = dvdy, S33 = dwdz;
  const R S12 = h * (dudy + dvdx), S13 = h * (dudz + dwdx), S23 = h * (dvdz + dwdy);
  const R mag = sqrt(R(2) * (S11 * S11 + S22 * S22 + S33 * S33 + R(2) * (S12 * S12 + S13 * S13 + S23 * S23)));
  R v = R(0);
  const R phi = a.phase ? ((const R*)a.phase)[c] : R(1);
  if (!(mag < R(1e-3)) && !(fabs(phi) < R(0.9))) {
    const R loc = (R)a.cs2 * mag;
    v = loc < (R)a.nu_max ? loc : (R)a.nu_max;
  }
  nu[c] = v;
}

hipError_t launch_les_fd(bool dbl, const LesFdArgs& a, hipStream_t s) {
  const dim3 grid((a.N + 255) / 256), block(256);
  if (dbl) hipLaunchKernelGGL(k_les_fd<double>, grid, block, 0, s, a);
  else hipLaunchKernelGGL(k_les_fd<float>, grid, block, 0, s, a);
  return hipGetLastError();
}

}  // namespace pol
