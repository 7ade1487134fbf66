// Hand-written discrete adjoint of the fused stream-collide step.
//
// Forward step n, cell x (not solid):
//   g_q(x)   = f^n_q(x - e_q)            (or f^n_{opp q}(x) on a wall link,
//                                          or the clamped source at an open face)
//   f*(x)    = C(g(x); K(x), u_in)       (cell_model.h)
//   f^{n+1}(x) = f*(x)
// Objective J = sum over objective steps of sum over outlet cells of -rho u_z.
//
// The adjoint buffer stores mu^n_q(x) = dJ/dg_q(x) at the reader cell x (a
// local write).  The adjoint of the post-collision value f^{n+1}_p(y) is a
// reverse gather of mu^{n+1}: the readers of f^{n+1}_p(y) are the cell
// y + e_p (normal link), y itself through its bounce link opp(p), and at open
// faces the readers whose source was clamped onto y.  One kernel per reverse
// step therefore does: reverse-gather lambda = dJ/df^{n+1}(x) from mu^{n+1},
// re-gather g(x) from the stored forward state f^n, and write
// mu^n(x) = (dC/dg)^T lambda + dJ_n/dg.  Parameter gradients: dJ/dK(x)
// accumulates in FP64 per cell, dJ/du_in and J in FP64 block partials.
//
// There is no adjoint in the reference; the forward semantics it inverts are
// those of lbm_step.hip / cell_model.h.
#include "cell_model.h"
#include "compact_gather.h"
#include "engine.h"

namespace pol {

// reverse-gather flag word (adjoint only, built on the host with the geometry)
//   bit p (0..18): the normal reader y + e_p of population p exists and gathers
//   bit 31       : y lies on a non-periodic face where clamped readers may exist

// Weighted moments of the adjoint populations: every sum the collision VJP
// needs is a polynomial in (u, F) of these ten numbers,
//   M0 = sum_p w_p lam_p,  M1 = sum_p w_p lam_p e_p,  M2 = sum_p w_p lam_p e_p e_p,
// accumulated with additions only (e components are -1, 0, 1).
template <typename R>
struct LamMoments {
  R m0, m1x, m1y, m1z, xx, yy, zz, xy, xz, yz;
  __device__ void m2dot(R x, R y, R z, R& ox, R& oy, R& oz) const {
    ox = xx * x + xy * y + xz * z;
    oy = xy * x + yy * y + yz * z;
    oz = xz * x + yz * y + zz * z;
  }
};
// Pairwise over opposite directions (w_a = w_b, e_b = -e_a): the even
// moments take w (lam_a + lam_b), the odd ones w (lam_a - lam_b).
template <typename R>
__device__ inline LamMoments<R> lam_moments(const R (&lam)[Q]) {
  LamMoments<R> M{};
  M.m0 = lam[0] * W<R>(0);
  static_for<0, 9>([&](auto pc) {
    constexpr int p = PAIR_A(decltype(pc)::value);
    constexpr int o = OPP(p);
    const R ws = (lam[p] + lam[o]) * W<R>(p);
    const R wd = (lam[p] - lam[o]) * W<R>(p);
    M.m0 += ws;
    eacc<EX(p)>(M.m1x, wd);
    eacc<EY(p)>(M.m1y, wd);
    eacc<EZ(p)>(M.m1z, wd);
    eacc<EX(p) * EX(p)>(M.xx, ws);
    eacc<EY(p) * EY(p)>(M.yy, ws);
    eacc<EZ(p) * EZ(p)>(M.zz, ws);
    eacc<EX(p) * EY(p)>(M.xy, ws);
    eacc<EX(p) * EZ(p)>(M.xz, ws);
    eacc<EY(p) * EZ(p)>(M.yz, ws);
  });
  return M;
}

// FP32: the hardware reciprocal and square root (1 ulp): the adjoint is
// linearised about the recomputed state, where a last-bit difference is
// immaterial; FP64 keeps the correctly rounded operations.
template <typename R>
__device__ __forceinline__ R arcp(R x) {
  if constexpr (std::is_same_v<R, float>) return __builtin_amdgcn_rcpf(x);
  else return R(1) / x;
}
template <typename R>
__device__ __forceinline__ R asqrt(R x) {
  if constexpr (std::is_same_v<R, float>) return __builtin_amdgcn_sqrtf(x);
  else return sqrt(x);
}
// sum_p lam_p d feq_p(rho, u)/du = rho (3 M1 + 9 M2 u - 3 u M0)
template <typename R>
__device__ inline void eq_grad_u(const LamMoments<R>& M, R rho, R ux, R uy, R uz, R& ax, R& ay, R& az) {
  R tx, ty, tz;
  M.m2dot(ux, uy, uz, tx, ty, tz);
  ax = rho * (R(3) * M.m1x + R(9) * tx - R(3) * ux * M.m0);
  ay = rho * (R(3) * M.m1y + R(9) * ty - R(3) * uy * M.m0);
  az = rho * (R(3) * M.m1z + R(9) * tz - R(3) * uz * M.m0);
}

// The bounce bits of a flag word moved to the opposite directions: bit p of
// the result is bit opp(p) of fl (bits 1..18; the pairs (1,2) (3,4) (5,6)
// (8,9) (12,13) (16,17) are one bit apart, (7,10) (11,14) (15,18) three).
__device__ __forceinline__ uint32_t opposite_links(uint32_t fl) {
  constexpr uint32_t LO1 = (1u << 1) | (1u << 3) | (1u << 5) | (1u << 8) | (1u << 12) | (1u << 16);
  constexpr uint32_t LO3 = (1u << 7) | (1u << 11) | (1u << 15);
  return ((fl & LO1) << 1) | ((fl >> 1) & LO1) | ((fl & LO3) << 3) | ((fl >> 3) & LO3);
}
static_assert(OPP(1) == 2 && OPP(3) == 4 && OPP(5) == 6 && OPP(8) == 9 && OPP(12) == 13 && OPP(16) == 17 &&
                  OPP(7) == 10 && OPP(11) == 14 && OPP(15) == 18,
              "opposite_links assumes the config/core.py direction order");

struct AdjGeom {
  uint32_t x, y, z;
  bool px, py, pz;
};

template <typename S, typename R, bool A32, bool SPARSE, bool XF = true>
__global__ void __launch_bounds__(BLOCK) k_adjoint(const AdjointArgs A, const uint32_t* __restrict__ rflags) {
  using Cd = Codec<S>;
  using AD = Addr<A32>;
  using AI = typename AD::idx;
  const StepArgs& a = A.fw;
  const uint32_t lblk = logical_block(a.per_xcd);
  if (lblk >= a.nblk) return;
  uint32_t cell, x, y, z;
  int rb[9];  // compact layout: slot bases of the nine neighbour rows
  int sx = 0;  // compact layout: the lane's unwrapped x (its slot is rb[4] + sx)
  bool in_run = false;
  bool plain = false;  // compact layout: every flag word of the piece is 0
  bool inside;
  if constexpr (SPARSE) {
    inside = item_record(lblk * BLOCK + threadIdx.x, a.n_items, a.chunks, a.records, a.cshift, a.N, a.dnx, a.nx,
                         cell, y, z, x, rb, sx, in_run, plain);
  } else {
    uint32_t yz;
    inside = item_cell(lblk * BLOCK + threadIdx.x, a.n_items, a.chunks, a.cshift, a.N, a.dnx, a.nx, cell, yz, x);
    z = fdiv(yz, a.dny);
    y = yz - z * a.ny;
  }
  const uint32_t N = a.N;  // dense cells
  const uint32_t P = a.P;  // population plane stride
  uint32_t own = cell;     // population slot of this cell
  if constexpr (SPARSE) own = slot_clamp(rb[4] + sx, P);
  const uint32_t fl = inside ? (plain ? 0u : a.flags[cell]) : (POL_KIND_SOLID << POL_KIND_SHIFT);
  const uint32_t rf = inside ? rflags[cell] : 0u;  // loaded beside fl: one round trip for both
  const uint32_t kind = (fl >> POL_KIND_SHIFT) & POL_KIND_MASK;
  const bool active = inside && kind != POL_KIND_SOLID;

  R part[4] = {R(0), R(0), R(0), R(0)};  // this cell's dJ/du_in (3) and J; summed in FP64
  const bool px = a.options & POL_PERIODIC_X, py = a.options & POL_PERIODIC_Y, pz = a.options & POL_PERIODIC_Z;
  const int nx = (int)a.nx, ny = (int)a.ny, nz = (int)a.nz, nxy = (int)a.nxy;
  const int oxm = x == 0 ? (px ? nx - 1 : 0) : -1;
  const int oxp = x == a.nx - 1 ? (px ? -(nx - 1) : 0) : 1;
  const int oym = y == 0 ? (py ? (ny - 1) * nx : 0) : -nx;
  const int oyp = y == a.ny - 1 ? (py ? -(ny - 1) * nx : 0) : nx;
  const int ozm = z == 0 ? (pz ? (nz - 1) * nxy : 0) : -nxy;
  const int ozp = z == a.nz - 1 ? (pz ? -(nz - 1) * nxy : 0) : nxy;

  // Every lane's output has the form
  //   mu_q = keep lam_q + C + e_q.h + L1 e_q.P.e_q
  // (fluid, porous and outlet cells set the scalars; solid lanes of a launched
  // chunk and inlet cells, whose adjoint no reader gathers, keep them zero),
  // so one shared loop writes the 19 values for every lane kind.
  R lam[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) lam[q] = R(0);
  R keep = 0, C = 0, hx = 0, hy = 0, hz = 0, L1 = 0;
  R Pxx = 0, Pyy = 0, Pzz = 0, Pxy2 = 0, Pxz2 = 0, Pyz2 = 0;
  if (active) {
    const R* __restrict__ mu_in = (const R*)A.lam_in;
    // ---- 1. lambda_p = dJ/df^{n+1}_p(x), reverse gather of mu^{n+1} ----
    if constexpr (SPARSE) {
      // one load per direction from the selected reader (compact_gather.h);
      // bit p of `readers`: f_p(x) has a reader (normal link or x's own bounce
      // link opp p), else lambda_p = 0
      const uint32_t readers = rf | (kind != POL_KIND_INLET ? opposite_links(fl) : 0u);
      compact_readers<R, A32, XF>(mu_in, P, rb, x, a.nx, px, own, rf, [&](auto pc, R v) {
        constexpr int p = decltype(pc)::value;
        lam[p] = (readers & (1u << p)) ? v : R(0);
      });
    } else static_for<0, Q>([&](auto qc) {
      constexpr int p = decltype(qc)::value;
      // normal reader at x + e_p
      const int off = (EX(p) > 0 ? oxp : (EX(p) < 0 ? oxm : 0)) + (EY(p) > 0 ? oyp : (EY(p) < 0 ? oym : 0)) +
                      (EZ(p) > 0 ? ozp : (EZ(p) < 0 ? ozm : 0));
      // The reader of f_p(x) is either x + e_p (normal link, rev bit p) or x
      // itself through its bounce link opp(p) -- never both: the bounce means
      // x + e_p is a wall or outside the box, which has no normal reader.  One
      // unconditional load from a select of the two addresses (own cell when
      // neither applies, then zeroed) keeps the gather free of branches.
      const bool nb = rf & (1u << p);
      bool bb = false;
      if constexpr (p != 0) bb = kind != POL_KIND_INLET && (fl & (1u << OPP(p)));
      const uint32_t rslot = (uint32_t)((int)cell + off);
      const AI src = nb ? (AI)p * P + rslot : (AI)OPP(p) * P + own;
      const R v = *AD::at(mu_in, src);
      lam[p] = (nb || bb) ? v : R(0);
    });
    // ---- 1b. re-gather the forward cell from the stored state ----
    // Issued right behind the reverse gather for every active lane (inlet
    // lanes ignore the values), so the two gathers share one memory round
    // trip instead of following each other.
    R dv[Q];  // stored deviations D_q = 36 (g_q - w_q) of the gathered cell
    {
      const S* __restrict__ fin = (const S*)a.fin;
      if constexpr (SPARSE) {
        compact_sources<S, A32, XF>(fin, P, rb, x, a.nx, px, own, fl,
                                    [&](auto qc, R v) { dv[decltype(qc)::value] = v; });
      } else {
        static_for<0, Q>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          const int off = (EX(q) > 0 ? oxm : (EX(q) < 0 ? oxp : 0)) + (EY(q) > 0 ? oym : (EY(q) < 0 ? oyp : 0)) +
                          (EZ(q) > 0 ? ozm : (EZ(q) < 0 ? ozp : 0));
          AI src = (AI)q * P + (uint32_t)((int)cell + off);
          if constexpr (q != 0) {
            if (fl & (1u << q)) src = (AI)OPP(q) * P + own;
          }
          dv[q] = Cd::load(AD::at(fin, src));
        });
      }
    }
    if (rf & POL_RFLAG_CLAMP) {
      // readers whose out-of-domain source was clamped onto this cell
      const int c[3] = {(int)x, (int)y, (int)z};
      const int n3[3] = {nx, ny, nz};
      const bool per[3] = {px, py, pz};
      static_for<1, Q>([&](auto qc) {
        constexpr int p = decltype(qc)::value;
        const int e[3] = {EX(p), EY(p), EZ(p)};
        int clampable = 0;
        for (int d = 0; d < 3; ++d)
          if (!per[d] && ((c[d] == 0 && e[d] == 1) || (c[d] == n3[d] - 1 && e[d] == -1))) clampable |= 1 << d;
        for (int sub = clampable; sub; sub = (sub - 1) & clampable) {
          int r[3];
          bool ok = true;
          for (int d = 0; d < 3; ++d) {
            if (sub & (1 << d)) {
              r[d] = c[d];
            } else {
              r[d] = c[d] + e[d];
              if (r[d] < 0 || r[d] >= n3[d]) {
                if (per[d]) r[d] += r[d] < 0 ? n3[d] : -n3[d];  // one period away at most
                else ok = false;
              }
            }
          }
          if (!ok) continue;
          const uint32_t rc = (uint32_t)(r[2] * nxy + r[1] * nx + r[0]);
          const uint32_t rfl = a.flags[rc];
          const uint32_t rk = (rfl >> POL_KIND_SHIFT) & POL_KIND_MASK;
          if (rk == POL_KIND_SOLID || rk == POL_KIND_INLET || (rfl & (1u << p))) continue;
          uint32_t rs = rc;
          if constexpr (SPARSE) {
            // the reader's row is one of the record's nine rows
            const int ra = (sub & 2) ? 1 : 1 + e[1], rbz = (sub & 4) ? 1 : 1 + e[2];
            rs = slot_clamp(rb[ra + 3 * rbz] + r[0], P);
          }
          lam[p] += *AD::at(mu_in, (AI)p * P + rs);
        }
      });
    }

    const LamMoments<R> M = lam_moments<R>(lam);

    // ---- 2. recompute the forward cell state ----
    const R G0 = (R)a.G[0], G1 = (R)a.G[1], G2 = (R)a.G[2];
    if (kind == POL_KIND_INLET) {
      const R ri = (R)a.inlet_rho, ux = (R)a.inlet_u[0], uy = (R)a.inlet_u[1], uz = (R)a.inlet_u[2];
      R gx, gy, gz;
      eq_grad_u<R>(M, ri, ux, uy, uz, gx, gy, gz);
      part[0] = gx;
      part[1] = gy;
      part[2] = gz;
      // nothing gathered: mu at inlet cells is never read (left zero)
    } else {
      // reduce the gathered deviations to what the adjoint needs: moments,
      // the second moment (LES stress) and sum_q lam_q D_q
      // pairwise: even moments from d_a + d_b, odd ones from d_a - d_b
      R s36 = dv[0], mx = 0, my = 0, mz = 0, lamD = lam[0] * dv[0];
      R pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
      static_for<0, 9>([&](auto pc) {
        constexpr int q = PAIR_A(decltype(pc)::value);
        constexpr int o = OPP(q);
        const R ds = dv[q] + dv[o], dd = dv[q] - dv[o];
        s36 += ds;
        lamD += lam[q] * dv[q] + lam[o] * dv[o];
        eacc<EX(q)>(mx, dd);
        eacc<EY(q)>(my, dd);
        eacc<EZ(q)>(mz, dd);
        eacc<EX(q) * EX(q)>(pxx, ds);
        eacc<EY(q) * EY(q)>(pyy, ds);
        eacc<EZ(q) * EZ(q)>(pzz, ds);
        eacc<EX(q) * EY(q)>(pxy, ds);
        eacc<EX(q) * EZ(q)>(pxz, ds);
        eacc<EY(q) * EZ(q)>(pyz, ds);
      });
      // the linearisation point of the adjoint: a rounded 1/36 is immaterial
      const R i36 = R(1.0 / 36.0);
      const R rd = s36 * i36, rho = R(1) + rd;
      mx *= i36;
      my *= i36;
      mz *= i36;
      const R inv_rho = arcp(rho);
      R F0x = G0, F0y = G1, F0z = G2;
      if (a.options & POL_FORCE_FIELD) {
        const R* Ff = (const R*)a.force;
        F0x += Ff[cell];
        F0y += Ff[(size_t)N + cell];
        F0z += Ff[2 * (size_t)N + cell];
      }
      const R vx = (mx + R(0.5) * F0x) * inv_rho, vy = (my + R(0.5) * F0y) * inv_rho,
              vz = (mz + R(0.5) * F0z) * inv_rho;
      if (kind == POL_KIND_OUTLET) {
        const R ro = (R)a.outlet_rho;
        R ax, ay, az;
        eq_grad_u<R>(M, ro, vx, vy, vz, ax, ay, az);
        // mu_q = a.(e_q - v)/rho - [objective] e_qz
        hx = ax * inv_rho;
        hy = ay * inv_rho;
        hz = az * inv_rho;
        C = -(hx * vx + hy * vy + hz * vz);
        if (A.objective) hz -= R(1);
        if (A.objective) part[3] = -(mz + R(0.5) * F0z);
      } else {
        const bool porous = fl & POL_FLAG_POROUS;
        const bool les = (a.options & POL_LES) && !(fl & POL_FLAG_NO_LES) && !porous;
        const R nu0 = (R)a.nu0, tau0 = (R)a.tau0;
        R K = (R)a.K;
        if (porous && (a.options & POL_PERM_FIELD)) K = ((const R*)a.perm)[cell];
        // velocity and total force
        R ux = vx, uy = vy, uz = vz, Fx = F0x, Fy = F0y, Fz = F0z;
        R dk = 0, fk = 0, umag = 0, sK = 0, c0 = 0, c1 = 0, vmag = 0, rr = 0, D = 0;
        if (porous) {
          sK = asqrt(K);
          const R iK = arcp(K), isK = arcp(sK);
          c0 = R(1) + R(0.5) * nu0 * iK;
          c1 = R(0.5) * (R)a.CF * isK;
          vmag = asqrt(vx * vx + vy * vy + vz * vz);
          rr = asqrt(c0 * c0 + R(4) * c1 * vmag);
          D = c0 + rr;
          const R s = R(2) * arcp(D);
          ux = vx * s;
          uy = vy * s;
          uz = vz * s;
          dk = nu0 * iK;
          fk = (R)a.CF * isK;
          umag = asqrt(ux * ux + uy * uy + uz * uz);
          const R cc = rho * (dk + fk * umag);
          Fx -= cc * ux;
          Fy -= cc * uy;
          Fz -= cc * uz;
        }
        const R usq = ux * ux + uy * uy + uz * uz;
        R tau = tau0, Qn = 0, sroot = 0;
        bool les_active = false;
        if (les) {
          // Pi_ab = sum_q e_a e_b (g_q - feq_q) = sum_q e_a e_b D_q / 36 - rd/3 d_ab - rho u_a u_b
          const R r3 = rd * R(1.0 / 3.0);
          pxx = pxx * i36 - r3 - rho * ux * ux;
          pyy = pyy * i36 - r3 - rho * uy * uy;
          pzz = pzz * i36 - r3 - rho * uz * uz;
          pxy = pxy * i36 - rho * ux * uy;
          pxz = pxz * i36 - rho * ux * uz;
          pyz = pyz * i36 - rho * uy * uz;
          Qn = asqrt(pxx * pxx + pyy * pyy + pzz * pzz + R(2) * (pxy * pxy + pxz * pxz + pyz * pyz));
          sroot = asqrt(tau0 * tau0 + (R)a.les_c * Qn * inv_rho);
          tau = R(0.5) * (tau0 + sroot);
          const R tmax = (R)a.tau_max;
          if (tau < tmax) {
            les_active = Qn > R(0);
          } else {
            tau = tmax;
          }
        }
        const R om = arcp(tau);
        const R cf = R(1) - R(0.5) * om;
        // ---- 3. adjoint of the collision, from the lambda moments ----
        //   sum lam feq / rho  = (M0 + rd M0 + rho (3 M1.u + 4.5 u.M2.u - 1.5 u.u M0)) / rho
        //   sum lam dfeq/du    = rho (3 M1 + 9 M2 u - 3 u M0)                  (Au)
        //   sum lam dS/du      = cf (9 M2 F - 3 F M0)                          (Bu)
        //   sum lam dS/dF      = cf (3 M1 - 3 u M0 + 9 M2 u)                   (CF3)
        //   sum lam df*/domega = sum lam (feq - g) - (3 M1.F - 3 u.F M0 + 9 u.M2.F) / 2
        const R uF = ux * Fx + uy * Fy + uz * Fz;
        R M2u[3], M2F[3];
        M.m2dot(ux, uy, uz, M2u[0], M2u[1], M2u[2]);
        M.m2dot(Fx, Fy, Fz, M2F[0], M2F[1], M2F[2]);
        const R M1u = M.m1x * ux + M.m1y * uy + M.m1z * uz;
        const R M1F = M.m1x * Fx + M.m1y * Fy + M.m1z * Fz;
        const R uM2u = ux * M2u[0] + uy * M2u[1] + uz * M2u[2];
        const R uM2F = ux * M2F[0] + uy * M2F[1] + uz * M2F[2];
        const R lam_fdev = rd * M.m0 + rho * (R(3) * M1u + R(4.5) * uM2u - R(1.5) * usq * M.m0);
        // the -g part: sum lam (w + D/36), less the w_p common to feq and g
        const R domega = lam_fdev - lamD * i36 - R(0.5) * (R(3) * M1F - R(3) * uF * M.m0 + R(9) * uM2F);
        const R Arho = (lam_fdev + M.m0) * inv_rho;
        const R u3[3] = {ux, uy, uz}, F3[3] = {Fx, Fy, Fz}, M1[3] = {M.m1x, M.m1y, M.m1z};
        R Au[3], Bu[3], CF3[3];
        for (int d = 0; d < 3; ++d) {
          Au[d] = rho * (R(3) * M1[d] + R(9) * M2u[d] - R(3) * u3[d] * M.m0);
          Bu[d] = cf * (R(-3) * F3[d] * M.m0 + R(9) * M2F[d]);
          // dS/dF = cf w (3 (e - u) + 9 (e.u) e)
          CF3[d] = cf * (R(3) * M1[d] - R(3) * u3[d] * M.m0 + R(9) * M2u[d]);
        }
        keep = R(1) - om;
        if (!porous) {
          // mu_q = (1-om) lam_q + om Arho + gu.(e_q - u)/rho
          //        + Lam ((w.P.w - trP/3)/Q - Q/rho),  w = e_q - u   (LES)
          // with w.P.w = e.P.e - 2 e.(P u) + u.P.u, so per direction only
          // additions remain: mu_q = keep lam_q + C + e.h + L1 e.P.e
          const R gx = (om * Au[0] + Bu[0]) * inv_rho, gy = (om * Au[1] + Bu[1]) * inv_rho,
                  gz = (om * Au[2] + Bu[2]) * inv_rho;
          C = om * Arho - (gx * ux + gy * uy + gz * uz);
          hx = gx;
          hy = gy;
          hz = gz;
          if (les_active) {
            const R Lam = domega * (-om * om) * (R)a.les_c * arcp(R(4) * sroot * rho);
            const R iQn = arcp(Qn);
            const R trP3 = (pxx + pyy + pzz) * R(1.0 / 3.0);
            const R Pux = pxx * ux + pxy * uy + pxz * uz, Puy = pxy * ux + pyy * uy + pyz * uz,
                    Puz = pxz * ux + pyz * uy + pzz * uz;
            const R uPu = ux * Pux + uy * Puy + uz * Puz;
            L1 = Lam * iQn;
            C += Lam * ((uPu - trP3) * iQn - Qn * inv_rho);
            hx -= R(2) * L1 * Pux;
            hy -= R(2) * L1 * Puy;
            hz -= R(2) * L1 * Puz;
            Pxx = pxx;
            Pyy = pyy;
            Pzz = pzz;
            Pxy2 = R(2) * pxy;
            Pxz2 = R(2) * pxz;
            Pyz2 = R(2) * pyz;
          }
        } else {
          // F = F0 - rho (dk + fk |u|) u ; u = v s(|v|)
          const R cdot = CF3[0] * ux + CF3[1] * uy + CF3[2] * uz;
          const R cc = dk + fk * umag;
          R gu[3];
          const R fcu = umag > R(0) ? rho * fk * cdot * arcp(umag) : R(0);
          for (int d = 0; d < 3; ++d) gu[d] = om * Au[d] + Bu[d] - rho * cc * CF3[d] - fcu * u3[d];
          const R grho = om * Arho - cc * cdot;
          // v-bar = (du/dv)^T gu,  du/dv = s I + (ds/d|v|) v v^T/|v|
          const R iD = arcp(D);
          const R s = R(2) * iD;
          const R dsdv = -R(4) * c1 * arcp(rr) * iD * iD;
          const R vg = vx * gu[0] + vy * gu[1] + vz * gu[2];
          R vb[3] = {s * gu[0], s * gu[1], s * gu[2]};
          if (vmag > R(0)) {
            const R t = dsdv * vg * arcp(vmag);
            vb[0] += t * vx;
            vb[1] += t * vy;
            vb[2] += t * vz;
          }
          // mu_q = (1-om) lam_q + grho + vb.(e_q - v)/rho
          hx = vb[0] * inv_rho;
          hy = vb[1] * inv_rho;
          hz = vb[2] * inv_rho;
          C = grho - (hx * vx + hy * vy + hz * vz);
          // dJ/dK: via u (c0, c1) and via F (dk, fk)
          const R iK2 = arcp(K * K), iK32 = arcp(K * sK);
          const R dc0 = R(-0.5) * nu0 * iK2;
          const R dc1 = R(-0.25) * (R)a.CF * iK32;
          const R dr = (c0 * dc0 + R(2) * vmag * dc1) * arcp(rr);
          const R dsdK = R(-2) * iD * iD * (dc0 + dr);
          const R du_dK = vg * dsdK;  // gu . (v ds/dK)
          const R ddk = -nu0 * iK2, dfk = R(-0.5) * (R)a.CF * iK32;
          const R dF = -rho * (ddk + dfk * umag) * cdot;
          if (A.grad_perm) A.grad_perm[cell] += (double)(du_dK + dF);
        }
      }
    }
  }
  {
    // One store loop for every lane kind.  Solid lanes of a launched chunk
    // store too (zeros), so every 128-byte output line is written whole;
    // dense launches do so only for 32-lane groups that hold a non-solid cell;
    // the compact layout stores exactly the slots of the stored runs.
    const uint64_t any = __ballot(active);
    const uint64_t group = 0xffffffffull << (threadIdx.x & 32);
    const bool store = SPARSE ? in_run : (inside && (active || a.chunks != nullptr || (any & group) != 0));
    if (store) {
      // pairwise: e.P.e is shared by q and opp q, e.h changes sign
      R* __restrict__ mu_out = (R*)A.lam_out;
      *AD::at(mu_out, own) = keep * lam[0] + C;
      static_for<0, 9>([&](auto pc) {
        constexpr int q = PAIR_A(decltype(pc)::value);
        constexpr int o = OPP(q);
        R ePe = R(0);
        eacc<EX(q) * EX(q)>(ePe, Pxx);
        eacc<EY(q) * EY(q)>(ePe, Pyy);
        eacc<EZ(q) * EZ(q)>(ePe, Pzz);
        eacc<EX(q) * EY(q)>(ePe, Pxy2);
        eacc<EX(q) * EZ(q)>(ePe, Pxz2);
        eacc<EY(q) * EZ(q)>(ePe, Pyz2);
        const R base = C + L1 * ePe;
        const R eh = edot<EX(q), EY(q), EZ(q)>(hx, hy, hz);
        *AD::at(mu_out + (AI)q * P, own) = keep * lam[q] + (base + eh);
        *AD::at(mu_out + (AI)o * P, own) = keep * lam[o] + (base - eh);
      });
    }
  }
  // block partials of dJ/du_in and J (FP64), added to the block's slot: the
  // slots hold the sums over every step since pol_adjoint_reset and are
  // folded once, in block order, by k_adjoint_final when collected
  // (only inlet and outlet cells contribute: other waves write zeros and skip
  // the FP64 shuffles)
  __shared__ double sh[BLOCK / 64][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const bool contributes = part[0] != R(0) || part[1] != R(0) || part[2] != R(0) || part[3] != R(0);
  if (__ballot(contributes)) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const double r = wave_sum((double)part[s]);
      if (lane == 0) sh[wid][s] = r;
    }
  } else if (lane == 0) {
#pragma unroll
    for (int s = 0; s < 4; ++s) sh[wid][s] = 0.0;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double r = 0.0;
    for (int w = 0; w < BLOCK / 64; ++w) r += sh[w][threadIdx.x];
    A.partial[(size_t)lblk * 4 + threadIdx.x] += r;  // this block's slot: summed over the steps
  }
}

__global__ void __launch_bounds__(256) k_adjoint_final(const double* __restrict__ partial, uint32_t nblk,
                                                       double* __restrict__ acc) {
  __shared__ double sh[4][4];
  double v[4] = {0, 0, 0, 0};
  for (uint32_t b = threadIdx.x; b < nblk; b += 256)
    for (int s = 0; s < 4; ++s) v[s] += partial[(size_t)b * 4 + s];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int s = 0; s < 4; ++s) {
    const double r = wave_sum(v[s]);
    if (lane == 0) sh[wid][s] = r;
  }
  __syncthreads();
  if (threadIdx.x < 4) acc[threadIdx.x] = sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
}

template <typename S, bool SPARSE>
static void launch_adj(const AdjointArgs& A, const uint32_t* rflags, dim3 grid, hipStream_t s) {
  const bool a32 = fits_addr32(Q, A.fw.P, sizeof(S));
  if constexpr (SPARSE) {
    if (!A.fw.xface) {  // compact layout with no run on an x face
      if (a32)
        hipLaunchKernelGGL((k_adjoint<S, S, true, true, false>), grid, dim3(BLOCK), 0, s, A, rflags);
      else
        hipLaunchKernelGGL((k_adjoint<S, S, false, true, false>), grid, dim3(BLOCK), 0, s, A, rflags);
      return;
    }
  }
  if (a32) {
    hipLaunchKernelGGL((k_adjoint<S, S, true, SPARSE>), grid, dim3(BLOCK), 0, s, A, rflags);
  } else {
    hipLaunchKernelGGL((k_adjoint<S, S, false, SPARSE>), grid, dim3(BLOCK), 0, s, A, rflags);
  }
}

hipError_t launch_adjoint(int storage, const AdjointArgs& A, const uint32_t* rflags, const Grid& g, hipStream_t s) {
  const dim3 grid(g.launch);
  const bool sparse = A.fw.records != nullptr;
  switch (storage) {
    case POL_F32:
      if (sparse) launch_adj<float, true>(A, rflags, grid, s);
      else launch_adj<float, false>(A, rflags, grid, s);
      break;
    case POL_F64:
      if (sparse) launch_adj<double, true>(A, rflags, grid, s);
      else launch_adj<double, false>(A, rflags, grid, s);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_adjoint_final(const double* partial, uint32_t nblk, double* acc, hipStream_t s) {
  hipLaunchKernelGGL(k_adjoint_final, dim3(1), dim3(256), 0, s, partial, nblk, acc);
  return hipGetLastError();
}

}  // namespace pol
