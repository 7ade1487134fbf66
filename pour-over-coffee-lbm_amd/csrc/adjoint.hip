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
#include "engine.h"

namespace pol {

// reverse-gather flag word (adjoint only, built on the host with the geometry)
//   bit p (0..18): the normal reader y + e_p of population p exists and gathers
//   bit 31       : y lies on a non-periodic face where clamped readers may exist

template <typename R>
__device__ inline void eq_grad_u(R rho, R ux, R uy, R uz, const R (&lam)[Q], R& ax, R& ay, R& az) {
  // sum_p lam_p d feq_p(rho,u)/du = rho sum_p lam_p w_p (3 e_p + 9 (e_p.u) e_p - 3 u)
  R sx = 0, sy = 0, sz = 0, s0 = 0;
  static_for<0, Q>([&](auto qc) {
    constexpr int p = decltype(qc)::value;
    const R lw = lam[p] * W<R>(p);
    const R eu = edot<EX(p), EY(p), EZ(p)>(ux, uy, uz);
    const R c = R(3) * lw + R(9) * lw * eu;
    eacc<EX(p)>(sx, c);
    eacc<EY(p)>(sy, c);
    eacc<EZ(p)>(sz, c);
    s0 += lw;
  });
  ax = rho * (sx - R(3) * ux * s0);
  ay = rho * (sy - R(3) * uy * s0);
  az = rho * (sz - R(3) * uz * s0);
}

struct AdjGeom {
  uint32_t x, y, z;
  bool px, py, pz;
};

template <typename S, typename R, bool A32>
__global__ void __launch_bounds__(BLOCK) k_adjoint(const AdjointArgs A, const uint32_t* __restrict__ rflags) {
  using Cd = Codec<S>;
  using AD = Addr<A32>;
  using AI = typename AD::idx;
  const StepArgs& a = A.fw;
  const uint32_t lblk = logical_block(a.per_xcd);
  if (lblk >= a.nblk) return;
  uint32_t cell, yz, x;
  const bool inside = item_cell(lblk * BLOCK + threadIdx.x, a.n_items, a.chunks, a.cshift, a.N, a.dnx, a.nx, cell, yz, x);
  const uint32_t N = a.N;
  const uint32_t z = fdiv(yz, a.dny);
  const uint32_t y = yz - z * a.ny;
  const uint32_t fl = inside ? a.flags[cell] : (POL_KIND_SOLID << POL_KIND_SHIFT);
  const uint32_t kind = (fl >> POL_KIND_SHIFT) & POL_KIND_MASK;
  const bool active = inside && kind != POL_KIND_SOLID;

  double part[4] = {0.0, 0.0, 0.0, 0.0};  // dJ/du_in (3), J
  const bool px = a.options & POL_PERIODIC_X, py = a.options & POL_PERIODIC_Y, pz = a.options & POL_PERIODIC_Z;
  const int nx = (int)a.nx, ny = (int)a.ny, nz = (int)a.nz, nxy = (int)a.nxy;
  const int oxm = x == 0 ? (px ? nx - 1 : 0) : -1;
  const int oxp = x == a.nx - 1 ? (px ? -(nx - 1) : 0) : 1;
  const int oym = y == 0 ? (py ? (ny - 1) * nx : 0) : -nx;
  const int oyp = y == a.ny - 1 ? (py ? -(ny - 1) * nx : 0) : nx;
  const int ozm = z == 0 ? (pz ? (nz - 1) * nxy : 0) : -nxy;
  const int ozp = z == a.nz - 1 ? (pz ? -(nz - 1) * nxy : 0) : nxy;

  if (active) {
    const R* __restrict__ mu_in = (const R*)A.lam_in;
    R* __restrict__ mu_out = (R*)A.lam_out;
    // ---- 1. lambda_p = dJ/df^{n+1}_p(x), reverse gather of mu^{n+1} ----
    R lam[Q];
    const uint32_t rf = rflags[cell];
    static_for<0, Q>([&](auto qc) {
      constexpr int p = decltype(qc)::value;
      // normal reader at x + e_p
      const int off = (EX(p) > 0 ? oxp : (EX(p) < 0 ? oxm : 0)) + (EY(p) > 0 ? oyp : (EY(p) < 0 ? oym : 0)) +
                      (EZ(p) > 0 ? ozp : (EZ(p) < 0 ? ozm : 0));
      R v = R(0);
      if (rf & (1u << p)) v = *AD::at(mu_in, (AI)p * N + (uint32_t)((int)cell + off));
      if constexpr (p != 0) {
        if (kind != POL_KIND_INLET && (fl & (1u << OPP(p)))) v += *AD::at(mu_in, (AI)OPP(p) * N + cell);
      }
      lam[p] = v;
    });
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
                if (per[d]) r[d] = (r[d] + n3[d]) % n3[d];
                else ok = false;
              }
            }
          }
          if (!ok) continue;
          const uint32_t rc = (uint32_t)(r[2] * nxy + r[1] * nx + r[0]);
          const uint32_t rfl = a.flags[rc];
          const uint32_t rk = (rfl >> POL_KIND_SHIFT) & POL_KIND_MASK;
          if (rk == POL_KIND_SOLID || rk == POL_KIND_INLET || (rfl & (1u << p))) continue;
          lam[p] += *AD::at(mu_in, (AI)p * N + rc);
        }
      });
    }

    // ---- 2. recompute the forward cell state ----
    const R G0 = (R)a.G[0], G1 = (R)a.G[1], G2 = (R)a.G[2];
    if (kind == POL_KIND_INLET) {
      const R ri = (R)a.inlet_rho, ux = (R)a.inlet_u[0], uy = (R)a.inlet_u[1], uz = (R)a.inlet_u[2];
      R gx, gy, gz;
      eq_grad_u<R>(ri, ux, uy, uz, lam, gx, gy, gz);
      part[0] = (double)gx;
      part[1] = (double)gy;
      part[2] = (double)gz;
      // nothing gathered: mu at inlet cells is never read
    } else {
      const S* __restrict__ fin = (const S*)a.fin;
      // Stream the gathered cell's stored deviations D_q = 36 (g_q - w_q)
      // once, reducing them on the fly to what the adjoint needs: moments,
      // the second moment (LES stress) and sum_q lam_q D_q.  No D array stays
      // live beside lam.
      R s36 = 0, mx = 0, my = 0, mz = 0, lamD = 0;
      R pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
      static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int off = (EX(q) > 0 ? oxm : (EX(q) < 0 ? oxp : 0)) + (EY(q) > 0 ? oym : (EY(q) < 0 ? oyp : 0)) +
                        (EZ(q) > 0 ? ozm : (EZ(q) < 0 ? ozp : 0));
        AI src = (AI)q * N + (uint32_t)((int)cell + off);
        if constexpr (q != 0) {
          if (fl & (1u << q)) src = (AI)OPP(q) * N + cell;
        }
        const R d = Cd::load(AD::at(fin, src));
        s36 += d;
        lamD += lam[q] * d;
        if constexpr (EX(q) > 0) mx += d;
        if constexpr (EX(q) < 0) mx -= d;
        if constexpr (EY(q) > 0) my += d;
        if constexpr (EY(q) < 0) my -= d;
        if constexpr (EZ(q) > 0) mz += d;
        if constexpr (EZ(q) < 0) mz -= d;
        if constexpr (EX(q) != 0) pxx += d;
        if constexpr (EY(q) != 0) pyy += d;
        if constexpr (EZ(q) != 0) pzz += d;
        if constexpr (EX(q) * EY(q) > 0) pxy += d;
        if constexpr (EX(q) * EY(q) < 0) pxy -= d;
        if constexpr (EX(q) * EZ(q) > 0) pxz += d;
        if constexpr (EX(q) * EZ(q) < 0) pxz -= d;
        if constexpr (EY(q) * EZ(q) > 0) pyz += d;
        if constexpr (EY(q) * EZ(q) < 0) pyz -= d;
      });
      // the linearisation point of the adjoint: a rounded 1/36 is immaterial
      const R i36 = R(1.0 / 36.0);
      const R rd = s36 * i36, rho = R(1) + rd;
      mx *= i36;
      my *= i36;
      mz *= i36;
      const R inv_rho = R(1) / rho;
      R F0x = G0, F0y = G1, F0z = G2;
      if (a.options & POL_FORCE_FIELD) {
        const R* Ff = (const R*)a.force;
        F0x += Ff[cell];
        F0y += Ff[(size_t)N + cell];
        F0z += Ff[2 * (size_t)N + cell];
      }
      const R vx = (mx + R(0.5) * F0x) * inv_rho, vy = (my + R(0.5) * F0y) * inv_rho,
              vz = (mz + R(0.5) * F0z) * inv_rho;
      auto put = [&](int q, R v) { *AD::at(mu_out, (AI)q * N + cell) = v; };
      if (kind == POL_KIND_OUTLET) {
        const R ro = (R)a.outlet_rho;
        R ax, ay, az;
        eq_grad_u<R>(ro, vx, vy, vz, lam, ax, ay, az);
        static_for<0, Q>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          R m = (ax * (R(EX(q)) - vx) + ay * (R(EY(q)) - vy) + az * (R(EZ(q)) - vz)) * inv_rho;
          if (A.objective) m -= R(EZ(q));
          put(q, m);
        });
        if (A.objective) part[3] = -(double)(mz + R(0.5) * F0z);
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
          sK = sqrt(K);
          c0 = R(1) + nu0 / (R(2) * K);
          c1 = (R)a.CF / (R(2) * sK);
          vmag = sqrt(vx * vx + vy * vy + vz * vz);
          rr = sqrt(c0 * c0 + R(4) * c1 * vmag);
          D = c0 + rr;
          const R s = R(2) / D;
          ux = vx * s;
          uy = vy * s;
          uz = vz * s;
          dk = nu0 / K;
          fk = (R)a.CF / sK;
          umag = sqrt(ux * ux + uy * uy + uz * uz);
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
          Qn = sqrt(pxx * pxx + pyy * pyy + pzz * pzz + R(2) * (pxy * pxy + pxz * pxz + pyz * pyz));
          sroot = sqrt(tau0 * tau0 + (R)a.les_c * Qn * inv_rho);
          tau = R(0.5) * (tau0 + sroot);
          const R tmax = (R)a.tau_max;
          if (tau < tmax) {
            les_active = Qn > R(0);
          } else {
            tau = tmax;
          }
        }
        const R om = R(1) / tau;
        const R cf = R(1) - R(0.5) * om;
        // ---- 3. adjoint of the collision ----
        // A_rho = sum lam fe / rho ; A_u = sum lam dfe/du ; B_u = sum lam dS/du ;
        // C_F = sum lam dS/dF ; dJ/domega
        R Arho = 0, Sw = 0, domega = 0;
        R lwe[3] = {0, 0, 0}, lweu[3] = {0, 0, 0}, lweF[3] = {0, 0, 0};
        const R uF = ux * Fx + uy * Fy + uz * Fz;
        static_for<0, Q>([&](auto qc) {
          constexpr int p = decltype(qc)::value;
          const R lw = lam[p] * W<R>(p);
          const R eu = edot<EX(p), EY(p), EZ(p)>(ux, uy, uz);
          const R eF = edot<EX(p), EY(p), EZ(p)>(Fx, Fy, Fz);
          // feq_p - w_p = w_p (rd + rho (3 e.u + 4.5 (e.u)^2 - 1.5 u.u))
          const R fdev = W<R>(p) * (rd + rho * (R(3) * eu + R(4.5) * eu * eu - R(1.5) * usq));
          Arho += lam[p] * fdev;
          Sw += lw;
          const R lweu_p = lw * eu, lweF_p = lw * eF;
          eacc<EX(p)>(lwe[0], lw);
          eacc<EY(p)>(lwe[1], lw);
          eacc<EZ(p)>(lwe[2], lw);
          eacc<EX(p)>(lweu[0], lweu_p);
          eacc<EY(p)>(lweu[1], lweu_p);
          eacc<EZ(p)>(lweu[2], lweu_p);
          eacc<EX(p)>(lweF[0], lweF_p);
          eacc<EY(p)>(lweF[1], lweF_p);
          eacc<EZ(p)>(lweF[2], lweF_p);
          // d f*_p / d omega = fe_p - g_p - S_p / (2 cf)
          domega += lam[p] * fdev - R(0.5) * lw * (R(3) * (eF - uF) + R(9) * eu * eF);
        });
        domega -= lamD * i36;  // the -g_p part of fe_p - g_p, less the w_p in both
        Arho = (Arho + Sw) * inv_rho;  // sum lam feq / rho with feq = w + fdev
        const R u3[3] = {ux, uy, uz}, F3[3] = {Fx, Fy, Fz};
        R Au[3], Bu[3], CF3[3];
        for (int d = 0; d < 3; ++d) {
          Au[d] = rho * (R(3) * lwe[d] + R(9) * lweu[d] - R(3) * u3[d] * Sw);
          Bu[d] = cf * (R(-3) * F3[d] * Sw + R(9) * lweF[d]);
          // dS/dF = cf w (3 (e - u) + 9 (e.u) e)
          CF3[d] = cf * (R(3) * lwe[d] - R(3) * u3[d] * Sw + R(9) * lweu[d]);
        }
        if (!porous) {
          R gu[3];
          for (int d = 0; d < 3; ++d) gu[d] = om * Au[d] + Bu[d];
          const R base = om * Arho;
          R Lam = 0, trP3 = 0, iQn = 0;
          if (les_active) {
            Lam = domega * (-om * om) * (R)a.les_c / (R(4) * sroot * rho);
            trP3 = (pxx + pyy + pzz) * R(1.0 / 3.0);
            iQn = R(1) / Qn;
          }
          static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const R wx = R(EX(q)) - ux, wy = R(EY(q)) - uy, wz = R(EZ(q)) - uz;
            R m = (R(1) - om) * lam[q] + base + (gu[0] * wx + gu[1] * wy + gu[2] * wz) * inv_rho;
            if (les_active) {
              const R wPw = wx * wx * pxx + wy * wy * pyy + wz * wz * pzz +
                            R(2) * (wx * wy * pxy + wx * wz * pxz + wy * wz * pyz);
              m += Lam * ((wPw - trP3) * iQn - Qn * inv_rho);
            }
            put(q, m);
          });
        } else {
          // F = F0 - rho (dk + fk |u|) u ; u = v s(|v|)
          const R cdot = CF3[0] * ux + CF3[1] * uy + CF3[2] * uz;
          const R cc = dk + fk * umag;
          R gu[3];
          for (int d = 0; d < 3; ++d) {
            gu[d] = om * Au[d] + Bu[d] - rho * cc * CF3[d];
            if (umag > R(0)) gu[d] -= rho * fk * cdot * u3[d] / umag;
          }
          const R grho = om * Arho - cc * cdot;
          // v-bar = (du/dv)^T gu,  du/dv = s I + (ds/d|v|) v v^T/|v|
          const R s = R(2) / D;
          const R dsdv = -R(4) * c1 / (rr * D * D);
          const R vg = vx * gu[0] + vy * gu[1] + vz * gu[2];
          R vb[3] = {s * gu[0], s * gu[1], s * gu[2]};
          if (vmag > R(0)) {
            const R t = dsdv * vg / vmag;
            vb[0] += t * vx;
            vb[1] += t * vy;
            vb[2] += t * vz;
          }
          static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            put(q, (R(1) - om) * lam[q] + grho +
                       (vb[0] * (R(EX(q)) - vx) + vb[1] * (R(EY(q)) - vy) + vb[2] * (R(EZ(q)) - vz)) * inv_rho);
          });
          // dJ/dK: via u (c0, c1) and via F (dk, fk)
          const R K2 = K * K, K32 = K * sK;
          const R dc0 = -nu0 / (R(2) * K2);
          const R dc1 = -(R)a.CF / (R(4) * K32);
          const R dr = (c0 * dc0 + R(2) * vmag * dc1) / rr;
          const R dsdK = -R(2) / (D * D) * (dc0 + dr);
          const R du_dK = vg * dsdK;  // gu . (v ds/dK)
          const R ddk = -nu0 / K2, dfk = -(R)a.CF / (R(2) * K32);
          const R dF = -rho * (ddk + dfk * umag) * cdot;
          if (A.grad_perm) A.grad_perm[cell] += (double)(du_dK + dF);
        }
      }
    }
  }
  {
    // Lanes that write no adjoint value (solid lanes of a launched chunk,
    // inlet cells, whose adjoint populations no reader gathers) store zeros,
    // so every 128-byte output line is written whole.  Dense launches do so
    // only for 32-lane groups that hold a written cell.
    const bool written = active && kind != POL_KIND_INLET;
    const uint64_t any = __ballot(written);
    const uint64_t group = 0xffffffffull << (threadIdx.x & 32);
    if (inside && !written && (a.chunks != nullptr || (any & group) != 0)) {
      R* __restrict__ mu_out = (R*)A.lam_out;
#pragma unroll
      for (int q = 0; q < Q; ++q) *AD::at(mu_out, (AI)q * N + cell) = R(0);
    }
  }
  // block partials of dJ/du_in and J (FP64)
  __shared__ double sh[BLOCK / 64][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const double r = wave_sum(part[s]);
    if (lane == 0) sh[wid][s] = r;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double r = 0.0;
    for (int w = 0; w < BLOCK / 64; ++w) r += sh[w][threadIdx.x];
    A.partial[(size_t)lblk * 4 + threadIdx.x] = r;
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
  if (threadIdx.x < 4) acc[threadIdx.x] += sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
}

hipError_t launch_adjoint(int storage, const AdjointArgs& A, const uint32_t* rflags, const Grid& g, hipStream_t s) {
  dim3 grid(g.launch), block(BLOCK);
  switch (storage) {
    case POL_F32:
      if (fits_addr32(Q, A.fw.N, sizeof(float)))
        hipLaunchKernelGGL((k_adjoint<float, float, true>), grid, block, 0, s, A, rflags);
      else
        hipLaunchKernelGGL((k_adjoint<float, float, false>), grid, block, 0, s, A, rflags);
      break;
    case POL_F64:
      if (fits_addr32(Q, A.fw.N, sizeof(double)))
        hipLaunchKernelGGL((k_adjoint<double, double, true>), grid, block, 0, s, A, rflags);
      else
        hipLaunchKernelGGL((k_adjoint<double, double, false>), grid, block, 0, s, A, rflags);
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
