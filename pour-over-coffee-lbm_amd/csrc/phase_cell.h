// Per-cell phase-field forms shared by the phase-field kernels and the fused
// pressure-force / surface-tension pass: cell decomposition, the interior
// test, the reference's vector norm, the stored normal of a neighbour formed
// on the fly from phi, and the CSF force of one cell
// (src/core/multiphase_3d.py:110-149, :365-381).  float32 with contraction
// off, correctly rounded division and square root.
#pragma once
#include "engine.h"

namespace pol {

__device__ inline void mp_decompose(const MultiphaseArgs& a, uint32_t i, uint32_t& x, uint32_t& y, uint32_t& z) {
  const uint32_t yz = fdiv(i, a.dnx);
  x = i - yz * a.nx;
  z = fdiv(yz, a.dny);
  y = yz - z * a.ny;
}

__device__ inline bool mp_interior(const MultiphaseArgs& a, uint32_t x, uint32_t y, uint32_t z) {
  return x >= 1 && x + 1 < a.nx && y >= 1 && y + 1 < a.ny && z >= 1 && z + 1 < a.nz;
}

// sqrt((x x + y y) + z z), the reference's vector norm
__device__ inline float norm3(float x, float y, float z) {
#pragma clang fp contract(off)
  return __builtin_sqrtf((x * x + y * y) + z * z);
}

// Component c of the stored normal at cell j = (x, y, z): grad phi / |grad phi|
// on interior cells (0 where |grad phi| <= 1e-10), 0 on face cells.
template <int COMP>
__device__ inline float normal_comp(const MultiphaseArgs& a, uint32_t j, uint32_t x, uint32_t y, uint32_t z) {
#pragma clang fp contract(off)
  if (!mp_interior(a, x, y, z)) return 0.0f;
  const float* __restrict__ phi = a.phi;
  const float gx = (phi[j + 1] - phi[j - 1]) * 0.5f;
  const float gy = (phi[j + a.nx] - phi[j - a.nx]) * 0.5f;
  const float gz = (phi[j + a.nxy] - phi[j - a.nxy]) * 0.5f;
  const float m = norm3(gx, gy, gz);
  if (!(m > 1e-10f)) return 0.0f;
  const float g = COMP == 0 ? gx : (COMP == 1 ? gy : gz);
  return g / m;
}

// CSF force of interior cell i (|phi| < 0.9 and |grad phi| > 1e-10, else 0):
// sigma kappa |grad phi| n with kappa = div n from the neighbours' normals.
struct CellForce {
  float gx, gy, gz, n0, n1, n2, kap, f0, f1, f2;
};
__device__ inline CellForce cell_force(const MultiphaseArgs& a, uint32_t i, uint32_t x, uint32_t y, uint32_t z,
                                       float p) {
#pragma clang fp contract(off)
  CellForce c{};
  const float* __restrict__ phi = a.phi;
  c.gx = (phi[i + 1] - phi[i - 1]) * 0.5f;
  c.gy = (phi[i + a.nx] - phi[i - a.nx]) * 0.5f;
  c.gz = (phi[i + a.nxy] - phi[i - a.nxy]) * 0.5f;
  const float gm = norm3(c.gx, c.gy, c.gz);
  if (gm > 1e-10f) {
    c.n0 = c.gx / gm;
    c.n1 = c.gy / gm;
    c.n2 = c.gz / gm;
  }
  if (norm3(c.n0, c.n1, c.n2) > 1e-10f) {
    const float dnx = (normal_comp<0>(a, i + 1, x + 1, y, z) - normal_comp<0>(a, i - 1, x - 1, y, z)) * 0.5f;
    const float dny = (normal_comp<1>(a, i + a.nx, x, y + 1, z) - normal_comp<1>(a, i - a.nx, x, y - 1, z)) * 0.5f;
    const float dnz = (normal_comp<2>(a, i + a.nxy, x, y, z + 1) - normal_comp<2>(a, i - a.nxy, x, y, z - 1)) * 0.5f;
    c.kap = (dnx + dny) + dnz;
  }
  if (fabsf(p) < 0.9f && gm > 1e-10f) {
    const float fm = (a.sigma * c.kap) * gm;
    c.f0 = fm * c.n0;
    c.f1 = fm * c.n1;
    c.f2 = fm * c.n2;
  }
  return c;
}

}  // namespace pol
