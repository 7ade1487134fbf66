// Per-cell physics of one D3Q19 step, shared by the forward stream-collide
// kernel and the adjoint kernel (which recomputes the forward cell state).
//
// Populations are carried as scaled deviations D_q = 36 (f_q - w_q) in every
// storage format and in registers.  36 w_q = c_q in {12, 2, 1} is exact, so
//   36 (rho - 1) = sum D,   36 rho u - 18 F = sum D e   (sum c e = 0),
//   36 (feq_q - w_q) = c_q [(rho - 1) + rho (3 e.u + 4.5 (e.u)^2 - 1.5 u.u)]
// is an exact rewrite of the raw-population algorithm in which no rounded
// weight appears: raw FP32 f ~ w loses ~1e-7 relative per rounding (rho
// drifts), and deviations with rounded weights w_q = 1/18, 1/36 give the
// equilibrium momentum a systematic (1 + 7.5e-9) factor per step.
//
// Reference semantics followed (reference file:line):
//   moments rho = sum f, u = (sum f e + F/2)/rho, guard rho > 1e-12
//       src/core/legacy/lbm_solver.py:488-535
//   equilibrium feq = w rho (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u.u)
//       src/core/lbm_algorithms.py:183-218, cpu_backend.py:122-125
//   BGK f* = f - (f - feq)/tau          src/core/backends/cpu_backend.py:129
//   Guo forcing, textbook form of docs/mathematical/mathematical_models.md:105-121:
//       F_q = (1 - 1/(2 tau)) w_q [3 (e_q - u).F + 9 (e_q.u)(e_q.F)]
//       (legacy/lbm_solver.py:747-764 uses 3 e.F + 9 (e.u)(u.F) instead)
//   tau_eff = tau + 3 nu_sgs, nu_sgs <= nu_sgs_max
//       legacy/lbm_solver.py:565-572, les_turbulence.py:369-375
//   Darcy-Forchheimer resistance F = -rho (nu/K) u - rho (C_F/sqrt K)|u| u
//       filter_paper.py:471-536 (physical-unit form), applied as a Guo force
//       with the implicit velocity of Guo & Zhao (Phys. Rev. E 66, 036304, 2002)
//   LES disabled in the filter layer     filter_paper.py:199-204
#pragma once
#include "lbm_common.h"

namespace pol {

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

// Physics constants in the arithmetic type R.
template <typename R>
struct Phys {
  R tau0;     // molecular relaxation time
  R omega0;   // 1 / tau0 (read by the 16-bit-storage collision only)
  R nu0;      // (tau0 - 1/2)/3
  R les_c;    // 18 sqrt(2) C_s^2
  R tau_max;  // tau0 + 3 nu_sgs_max
  R G[3];     // constant body-force density
  R K;        // uniform permeability
  R CF;       // Forchheimer coefficient
  R inlet_rho, inlet_u[3], outlet_rho;
};

// e . (x, y, z) for a lattice direction with components in {-1, 0, 1},
// from additions only: x * 0 cannot be folded under IEEE rules (NaN, signed
// zero), so a product with the component would keep every dead term alive.
template <int c, typename R>
__device__ __forceinline__ R signed_term(R v) {
  if constexpr (c > 0) return v;
  else return -v;
}
template <int ex, int ey, int ez, typename R>
__device__ __forceinline__ R edot(R x, R y, R z) {
  if constexpr (ex != 0 && ey != 0 && ez != 0) return signed_term<ex>(x) + signed_term<ey>(y) + signed_term<ez>(z);
  else if constexpr (ex != 0 && ey != 0) return signed_term<ex>(x) + signed_term<ey>(y);
  else if constexpr (ex != 0 && ez != 0) return signed_term<ex>(x) + signed_term<ez>(z);
  else if constexpr (ey != 0 && ez != 0) return signed_term<ey>(y) + signed_term<ez>(z);
  else if constexpr (ex != 0) return signed_term<ex>(x);
  else if constexpr (ey != 0) return signed_term<ey>(y);
  else if constexpr (ez != 0) return signed_term<ez>(z);
  else return R(0);
}
// acc += c v for a component c in {-1, 0, 1}
template <int c, typename R>
__device__ __forceinline__ void eacc(R& acc, R v) {
  if constexpr (c > 0) acc += v;
  else if constexpr (c < 0) acc -= v;
}

// 36 (feq_q - w_q) for density rho = 1 + rd (c = 36 w_q, exact)
template <typename R, int q>
__device__ inline R deq_q(R c, R rd, R rho, R ux, R uy, R uz, R usq) {
  const R eu = edot<EX(q), EY(q), EZ(q)>(ux, uy, uz);
  return c * (rd + rho * (R(3) * eu + R(4.5) * eu * eu - R(1.5) * usq));
}

// Moments of the collided state (for reductions / macro output).
template <typename R>
struct CellOut {
  R rd;  // rho - 1
  R ux, uy, uz;
};

// Opposite-direction pairs (a, opp a): e_opp = -e, w_opp = w.  The nine
// first members in config/core.py order; the rest population q = 0 is alone.
__host__ __device__ constexpr int PAIR_A(int p) {
  constexpr int t[9] = {1, 3, 5, 7, 8, 11, 12, 15, 16};
  return t[p];
}

// d = 36 (feq(rho, u) - w) written pairwise: for e_b = -e_a
//   36 (feq_a - w) = sym + anti, 36 (feq_b - w) = sym - anti,
//   sym = c (rd + rho (4.5 (e.u)^2 - 1.5 u.u)), anti = 3 c rho e.u
template <typename R>
__device__ inline void set_equilibrium(R (&d)[Q], R rd, R rho, R ux, R uy, R uz) {
  const R usq = ux * ux + uy * uy + uz * uz;
  const R base = rd - R(1.5) * rho * usq;
  d[0] = CW<R>(0) * base;
  static_for<0, 9>([&](auto pc) {
    constexpr int a = PAIR_A(decltype(pc)::value);
    constexpr int b = OPP(a);
    const R eu = edot<EX(a), EY(a), EZ(a)>(ux, uy, uz);
    const R sym = CW<R>(a) * (base + R(4.5) * rho * eu * eu);
    const R anti = R(3) * CW<R>(a) * rho * eu;
    d[a] = sym + anti;
    d[b] = sym - anti;
  });
}

// Collide one gathered cell in place: d[] in = pre-collision scaled
// deviations D = 36 (f - w), out = post-collision ones.  kind/porous from the
// cell flags; K = permeability, (fx,fy,fz) = extra force density; (bx,by,bz,B)
// = particle-drag field: the particles exert F_p = BV - B u on the cell.
//
// Drag-type forces are implicit in u: with Guo's rho u = m + F/2 and
// F = F0 + (BV - B u) - rho (nu/K + C_F |u| / sqrt K) u (porous cells only),
//   u (c0 + c1 |u|) = v,  v = (m + F0/2 + BV/2) / rho,
//   c0 = 1 + B / (2 rho) + nu / (2K),  c1 = C_F / (2 sqrt K),
// solved in closed form (Guo & Zhao 2002) -- stable for any K and any B.
//
// Evaluated pairwise over opposite directions: equilibrium and Guo source of
// a pair share their symmetric part and negate the antisymmetric one, and the
// non-equilibrium stress needs only the pair sums (e_b e_b = e_a e_a), which
// halves the arithmetic of a per-direction evaluation.
// LES: nu_fd >= 0 is a given eddy viscosity (the reference's finite-difference
// form, POL_LES_FD): tau = clamp(tau0 + 3 nu_fd, 0.55, 1.90) as the legacy
// collision clamps it (legacy/lbm_solver.py:565-572); nu_fd < 0 selects the
// local non-equilibrium-stress form below.
template <typename R>
__device__ __forceinline__ R tau_of_nu(R tau0, R nu_fd) {
  const R t = tau0 + R(3) * nu_fd;
  return t < R(0.55) ? R(0.55) : (t > R(1.90) ? R(1.90) : t);
}

template <typename R, bool FULL>
__device__ inline CellOut<R> collide_cell(R (&d)[Q], uint32_t kind, bool porous, bool les, R K, R fx, R fy, R fz,
                                          R bx, R by, R bz, R B, const Phys<R>& P, R nu_fd = R(-1)) {
  CellOut<R> o;
  if (FULL && kind == POL_KIND_INLET) {
    o.rd = P.inlet_rho - R(1);
    o.ux = P.inlet_u[0];
    o.uy = P.inlet_u[1];
    o.uz = P.inlet_u[2];
    set_equilibrium<R>(d, o.rd, P.inlet_rho, o.ux, o.uy, o.uz);
    return o;
  }
  // moments: 36 (rho - 1) = sum d, 36 m = sum_pairs (d_a - d_b) e_a
  R s36 = d[0], mx = R(0), my = R(0), mz = R(0);
  static_for<0, 9>([&](auto pc) {
    constexpr int a = PAIR_A(decltype(pc)::value);
    constexpr int b = OPP(a);
    s36 += d[a] + d[b];
    const R A = d[a] - d[b];
    if constexpr (EX(a) > 0) mx += A;
    if constexpr (EX(a) < 0) mx -= A;
    if constexpr (EY(a) > 0) my += A;
    if constexpr (EY(a) < 0) my -= A;
    if constexpr (EZ(a) > 0) mz += A;
    if constexpr (EZ(a) < 0) mz -= A;
  });
  const R rd = s36 / R(36);
  const R rho = R(1) + rd;
  const R rho36 = R(36) + s36;
  const R inv36 = rho36 > R(36e-12) ? R(1) / rho36 : R(0);  // 1 / (36 rho)
  R Fx = R(0), Fy = R(0), Fz = R(0);
  if constexpr (FULL) {
    Fx = P.G[0] + fx;
    Fy = P.G[1] + fy;
    Fz = P.G[2] + fz;
  }
  R ux, uy, uz;
  if (FULL && (porous || B > R(0))) {
    const R vx = (mx + R(18) * (Fx + bx)) * inv36;
    const R vy = (my + R(18) * (Fy + by)) * inv36;
    const R vz = (mz + R(18) * (Fz + bz)) * inv36;
    R c0 = R(1) + R(18) * B * inv36, c1 = R(0), dk = R(0), fk = R(0);
    if (porous) {
      const R sqK = sqrt(K);
      dk = P.nu0 / K;    // Darcy coefficient
      fk = P.CF / sqK;   // Forchheimer coefficient
      c0 += R(0.5) * dk;
      c1 = R(0.5) * fk;
    }
    const R vmag = sqrt(vx * vx + vy * vy + vz * vz);
    const R s = R(2) / (c0 + sqrt(c0 * c0 + R(4) * c1 * vmag));
    ux = vx * s;
    uy = vy * s;
    uz = vz * s;
    const R umag = sqrt(ux * ux + uy * uy + uz * uz);
    const R c = rho * (dk + fk * umag) + B;
    Fx += bx - c * ux;
    Fy += by - c * uy;
    Fz += bz - c * uz;
  } else {
    ux = (mx + R(18) * Fx) * inv36;
    uy = (my + R(18) * Fy) * inv36;
    uz = (mz + R(18) * Fz) * inv36;
  }
  o.rd = rd;
  o.ux = ux;
  o.uy = uy;
  o.uz = uz;
  if (FULL && kind == POL_KIND_OUTLET) {
    set_equilibrium<R>(d, P.outlet_rho - R(1), P.outlet_rho, ux, uy, uz);
    return o;
  }
  const R usq = ux * ux + uy * uy + uz * uz;
  const R base = rd - R(1.5) * rho * usq;  // shared part of every equilibrium
  const R d0eq = CW<R>(0) * base;
  R tau = P.tau0;
  if (FULL && les && !porous && nu_fd >= R(0)) {
    tau = tau_of_nu(P.tau0, nu_fd);
  } else if (FULL && les && !porous) {
    // 36 x non-equilibrium stress Pi_ab = sum_q e_a e_b (f_q - feq_q); each pair
    // contributes e_a e_a^T (n_a + n_b) with n_a + n_b = (d_a + d_b) - 2 sym
    R pxx = R(0), pyy = R(0), pzz = R(0), pxy = R(0), pxz = R(0), pyz = R(0);
    static_for<0, 9>([&](auto pc) {
      constexpr int a = PAIR_A(decltype(pc)::value);
      constexpr int b = OPP(a);
      const R eu = edot<EX(a), EY(a), EZ(a)>(ux, uy, uz);
      const R n = (d[a] + d[b]) - R(2) * CW<R>(a) * (base + R(4.5) * rho * eu * eu);
      if constexpr (EX(a) != 0) pxx += n;
      if constexpr (EY(a) != 0) pyy += n;
      if constexpr (EZ(a) != 0) pzz += n;
      if constexpr (EX(a) * EY(a) > 0) pxy += n;
      if constexpr (EX(a) * EY(a) < 0) pxy -= n;
      if constexpr (EX(a) * EZ(a) > 0) pxz += n;
      if constexpr (EX(a) * EZ(a) < 0) pxz -= n;
      if constexpr (EY(a) * EZ(a) > 0) pyz += n;
      if constexpr (EY(a) * EZ(a) < 0) pyz -= n;
    });
    const R q36 = sqrt(pxx * pxx + pyy * pyy + pzz * pzz + R(2) * (pxy * pxy + pxz * pxz + pyz * pyz));
    // tau = (tau0 + sqrt(tau0^2 + 18 sqrt2 Cs^2 Q / rho)) / 2 with Q / rho = q36 / (36 rho)
    tau = R(0.5) * (P.tau0 + sqrt(P.tau0 * P.tau0 + P.les_c * q36 * inv36));
    tau = tau < P.tau_max ? tau : P.tau_max;
  }
  const R omega = R(1) / tau;
  const R keep = R(1) - omega;
  // BGK + Guo: d* = (1 - w) d + w deq + S with
  //   S_q = (1 - w/2) c_q [3 (e_q - u).F + 9 (e_q.u)(e_q.F)]
  const R cf = FULL ? (R(1) - R(0.5) * omega) : R(0);
  const R uF = FULL ? (ux * Fx + uy * Fy + uz * Fz) : R(0);
  d[0] = keep * d[0] + omega * d0eq - (FULL ? cf * (R(3) * CW<R>(0) * uF) : R(0));
  static_for<0, 9>([&](auto pc) {
    constexpr int a = PAIR_A(decltype(pc)::value);
    constexpr int b = OPP(a);
    constexpr R c = CW<R>(a);
    const R eu = edot<EX(a), EY(a), EZ(a)>(ux, uy, uz);
    // omega multiplies the finished equilibrium terms: a rounded product such
    // as omega*3 would scale every cell's equilibrium momentum by the same
    // (1 + eps) and drift the flow systematically
    R sym = omega * (c * (base + R(4.5) * rho * eu * eu));
    R anti = omega * (R(3) * c * rho * eu);
    if constexpr (FULL) {
      const R eF = edot<EX(a), EY(a), EZ(a)>(Fx, Fy, Fz);
      sym += cf * (c * (R(9) * eu * eF - R(3) * uF));
      anti += cf * (R(3) * c * eF);
    }
    d[a] = keep * d[a] + sym + anti;
    d[b] = keep * d[b] + sym - anti;
  });
  return o;
}

// The same collision for 16-bit population storage.  Storing a deviation in
// FP16 (BF16) rounds it to 2^-11 (2^-8) relative, three to five orders above
// FP32 arithmetic rounding, so this form spends fewer instructions where the
// FP32 form spends them on exactness:
//   - 1 / (36 rho), 1 / tau, sqrt and 1/sqrt(K) use the hardware
//     approximations (v_rcp_f32, v_sqrt_f32, v_rsq_f32: 1 ulp);
//   - the non-equilibrium stress comes from the second moments of the pair
//     sums, 36 Pi_neq = sum_q D_q e_q e_q - (12 rd I + 36 rho u u), which is
//     the pairwise sum of collide_cell rewritten with the D3Q19 identities
//     sum_q c_q e_a e_b = 12 delta_ab, sum_q c_q (e.u)^2 e_a e_b =
//     4 (u.u delta_ab + 2 u_a u_b);
//   - BGK + Guo per pair from five per-cell coefficients:
//     sym = c (omega base - 3 cf u.F) + e.u (4.5 c omega rho e.u + 9 c cf e.F),
//     anti = 3 c omega rho e.u + 3 c cf e.F.
// Mass and momentum stay conserved to rounding (the identities above hold
// term by term).  Physics, branches and outputs are those of collide_cell.
// FP16 operands straight from the loaded 16-bit words (low half of a VGPR):
// v_fma_mix_f32 widens them exactly inside the FMA, so a + b, a - b and
// k a + c round once, exactly as the same operations on converted values --
// without the 19 separate conversions.
__device__ inline float mix_add(uint32_t a, uint32_t b) {
  float r;
  asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,1]" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ inline float mix_sub(uint32_t a, uint32_t b) {
  float r;
  asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel_hi:[1,0,1]" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ inline float mix_fma(float k, uint32_t a, float c) {
  float r;
  asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[0,1,0]" : "=v"(r) : "v"(k), "v"(a), "v"(c));
  return r;
}

// MIX: the populations 1..18 come as FP16 words in h (d[0] decoded); the
// pair sums, differences and the relaxation read them through mix_*.
template <bool FULL, bool MIX = false>
__device__ inline CellOut<float> collide_cell_lp(float (&d)[Q], uint32_t kind, bool porous, bool les, float K,
                                                 float fx, float fy, float fz, float bx, float by, float bz, float B,
                                                 const Phys<float>& P, const uint32_t (&h)[Q], float nu_fd = -1.0f) {
  using R = float;
  CellOut<R> o;
  if (FULL && kind == POL_KIND_INLET) {
    o.rd = P.inlet_rho - R(1);
    o.ux = P.inlet_u[0];
    o.uy = P.inlet_u[1];
    o.uz = P.inlet_u[2];
    set_equilibrium<R>(d, o.rd, P.inlet_rho, o.ux, o.uy, o.uz);
    return o;
  }
  R sp[9];  // pair sums d_a + d_b
  R s36 = d[0], mx = R(-0.0), my = R(-0.0), mz = R(-0.0);  // -0: the additive identity, folded
  static_for<0, 9>([&](auto pc) {
    constexpr int p = decltype(pc)::value;
    constexpr int a = PAIR_A(p);
    constexpr int b = OPP(a);
    R A;
    if constexpr (MIX) {
      sp[p] = mix_add(h[a], h[b]);
      A = mix_sub(h[a], h[b]);
    } else {
      sp[p] = d[a] + d[b];
      A = d[a] - d[b];
    }
    s36 += sp[p];
    eacc<EX(a)>(mx, A);
    eacc<EY(a)>(my, A);
    eacc<EZ(a)>(mz, A);
  });
  const R rd = s36 * R(1.0 / 36.0);
  const R rho = R(1) + rd;
  const R rho36 = R(36) + s36;
  const R inv36 = rho36 > R(36e-12) ? __builtin_amdgcn_rcpf(rho36) : R(0);  // 1 / (36 rho)
  R Fx = R(0), Fy = R(0), Fz = R(0);
  if constexpr (FULL) {
    Fx = P.G[0] + fx;
    Fy = P.G[1] + fy;
    Fz = P.G[2] + fz;
  }
  R ux, uy, uz;
  if (FULL && (porous || B > R(0))) {
    const R vx = (mx + R(18) * (Fx + bx)) * inv36;
    const R vy = (my + R(18) * (Fy + by)) * inv36;
    const R vz = (mz + R(18) * (Fz + bz)) * inv36;
    R c0 = R(1) + R(18) * B * inv36, c1 = R(0), dk = R(0), fk = R(0);
    if (porous) {
      dk = P.nu0 * __builtin_amdgcn_rcpf(K);
      fk = P.CF * __builtin_amdgcn_rsqf(K);
      c0 += R(0.5) * dk;
      c1 = R(0.5) * fk;
    }
    const R vmag = __builtin_amdgcn_sqrtf(vx * vx + vy * vy + vz * vz);
    const R s = R(2) * __builtin_amdgcn_rcpf(c0 + __builtin_amdgcn_sqrtf(c0 * c0 + R(4) * c1 * vmag));
    ux = vx * s;
    uy = vy * s;
    uz = vz * s;
    const R umag = __builtin_amdgcn_sqrtf(ux * ux + uy * uy + uz * uz);
    const R c = rho * (dk + fk * umag) + B;
    Fx += bx - c * ux;
    Fy += by - c * uy;
    Fz += bz - c * uz;
  } else {
    ux = (mx + R(18) * Fx) * inv36;
    uy = (my + R(18) * Fy) * inv36;
    uz = (mz + R(18) * Fz) * inv36;
  }
  o.rd = rd;
  o.ux = ux;
  o.uy = uy;
  o.uz = uz;
  if (FULL && kind == POL_KIND_OUTLET) {
    set_equilibrium<R>(d, P.outlet_rho - R(1), P.outlet_rho, ux, uy, uz);
    return o;
  }
  const R usq = ux * ux + uy * uy + uz * uz;
  const R base = rd - R(1.5) * rho * usq;
  R omega = P.omega0;
  if (FULL && les && !porous && nu_fd >= R(0)) {
    omega = __builtin_amdgcn_rcpf(tau_of_nu(P.tau0, nu_fd));
  } else if (FULL && les && !porous) {
    R sxx = R(-0.0), syy = R(-0.0), szz = R(-0.0), sxy = R(-0.0), sxz = R(-0.0), syz = R(-0.0);
    static_for<0, 9>([&](auto pc) {
      constexpr int p = decltype(pc)::value;
      constexpr int a = PAIR_A(p);
      if constexpr (EX(a) != 0) sxx += sp[p];
      if constexpr (EY(a) != 0) syy += sp[p];
      if constexpr (EZ(a) != 0) szz += sp[p];
      eacc<EX(a) * EY(a)>(sxy, sp[p]);
      eacc<EX(a) * EZ(a)>(sxz, sp[p]);
      eacc<EY(a) * EZ(a)>(syz, sp[p]);
    });
    const R r36 = R(36) * rho, i12 = R(12) * rd;
    const R pxx = sxx - (i12 + r36 * ux * ux), pyy = syy - (i12 + r36 * uy * uy), pzz = szz - (i12 + r36 * uz * uz);
    const R pxy = sxy - r36 * ux * uy, pxz = sxz - r36 * ux * uz, pyz = syz - r36 * uy * uz;
    const R q36 = __builtin_amdgcn_sqrtf(pxx * pxx + pyy * pyy + pzz * pzz + R(2) * (pxy * pxy + pxz * pxz + pyz * pyz));
    R tau = R(0.5) * (P.tau0 + __builtin_amdgcn_sqrtf(P.tau0 * P.tau0 + P.les_c * q36 * inv36));
    tau = tau < P.tau_max ? tau : P.tau_max;
    omega = __builtin_amdgcn_rcpf(tau);
  }
  const R keep = R(1) - omega;
  const R cf = FULL ? (R(1) - R(0.5) * omega) : R(0);
  const R uF = FULL ? (ux * Fx + uy * Fy + uz * Fz) : R(0);
  // per-cell coefficients for c = 1; c = 2 doubles them exactly
  const R S0 = omega * base - R(3) * cf * uF;
  const R S1 = R(4.5) * omega * rho, S2 = R(9) * cf;
  const R A1 = R(3) * omega * rho, A2 = R(3) * cf;
  d[0] = keep * d[0] + R(12) * S0;
  static_for<0, 9>([&](auto pc) {
    constexpr int a = PAIR_A(decltype(pc)::value);
    constexpr int b = OPP(a);
    constexpr R c = CW<R>(a);
    const R eu = edot<EX(a), EY(a), EZ(a)>(ux, uy, uz);
    R t = (c * S1) * eu, anti = (c * A1) * eu;
    if constexpr (FULL) {
      const R eF = edot<EX(a), EY(a), EZ(a)>(Fx, Fy, Fz);
      t += (c * S2) * eF;
      anti += (c * A2) * eF;
    }
    const R sym = eu * t + c * S0;
    if constexpr (MIX) {
      d[a] = mix_fma(keep, h[a], sym + anti);
      d[b] = mix_fma(keep, h[b], sym - anti);
    } else {
      d[a] = keep * d[a] + (sym + anti);
      d[b] = keep * d[b] + (sym - anti);
    }
  });
  return o;
}

}  // namespace pol
