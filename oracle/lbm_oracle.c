/*
 * lbm_oracle.c -- CPU restatement of the reference's D3Q19 hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library and its
 * Python host layer) may load or call this; it is used by tests/, by
 * __graft_entry__.smoke() as the reference of comparison, and by bench.py's cpu_baseline
 * leg.  Compiled twice from this one file: REAL=double (suffix _f64) and
 * REAL=float (suffix _f32); OpenMP over z-planes.
 *
 * Functions and the reference code they restate (reference file:line):
 *   oracle_step_*          one stream-collide step in the build's semantics:
 *                          pull streaming + halfway bounce-back
 *                          (legacy/lbm_solver.py:609-628 in push form), moments
 *                          with the Guo half-force (legacy/lbm_solver.py:519-534),
 *                          f_eq (lbm_algorithms.py:183-218), BGK (cpu_backend.py:129),
 *                          textbook Guo forcing (docs/mathematical/mathematical_models.md:105-121),
 *                          tau_eff = tau + 3 nu_sgs (legacy/lbm_solver.py:565-572),
 *                          Darcy-Forchheimer resistance (filter_paper.py:471-536),
 *                          statistics (visualizer.py:120-167, lbm_diagnostics.py:445-460)
 *   oracle_cpu_backend_*   CPUBackend.execute_collision_streaming as written
 *                          (cpu_backend.py:79-190): its own D3Q19 order
 *                          (:69-72), collision on non-solid cells with the
 *                          f >= 0 clamp (:91-138), pull streaming that keeps the
 *                          cell's own pre-collision value when the source is
 *                          outside the box (:140-166), destructive solid pass
 *                          with the hard-coded opposite table (:168-190)
 *   oracle_particle_index  coffee_particles.py:1159-1174 in float32
 *   oracle_particles_*     coffee_particles.py:1088-1154,1200-1212 (drag,
 *                          scatter into the implicit drag field or the
 *                          reference's explicit reaction field
 *                          :1146-1149,1049-1086, under-relaxation) and the
 *                          build's particle advance (replacing :641-720)
 *
 * Parity pins (tests/): closed forms (lattice identities, Taylor-Green decay,
 * Guo momentum law, porous terminal velocity, LES uniform shear, settling
 * velocity), invariants (sealed-box mass, streaming permutation), and the
 * reference's own known-answer positions for the particle index/gather.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/pourover_hip.h"

#ifndef REAL
#define REAL double
#endif
#ifndef SFX
#define SFX(n) n##_f64
#endif
typedef REAL real;

static const int EX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int EY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};

static int opp_of(int q) {
  for (int p = 0; p < 19; ++p)
    if (EX[p] == -EX[q] && EY[p] == -EY[q] && EZ[p] == -EZ[q]) return p;
  return -1;
}
static real wq(int q) { return q == 0 ? (real)1 / (real)3 : (q < 7 ? (real)1 / (real)18 : (real)1 / (real)36); }

static real feq(int q, real rho, real ux, real uy, real uz) {
  real eu = (real)EX[q] * ux + (real)EY[q] * uy + (real)EZ[q] * uz;
  real usq = ux * ux + uy * uy + uz * uz;
  return wq(q) * rho * ((real)1 + (real)3 * eu + (real)4.5 * eu * eu - (real)1.5 * usq);
}

typedef struct {
  real rho, ux, uy, uz;
} moments;

/* Collide one gathered cell in place (post-collision written back into g).
 * pd = particle-drag field values {sum w beta v (3), B = sum w beta} or NULL:
 * the particles exert BV - B u with u implicit. */
static moments collide(real *g, unsigned kind, int porous, int les, real K, const real F0[3], const real *pd,
                       const pol_desc *d, real nu_fd) {
  moments m;
  const real tau0 = (real)d->tau, nu0 = (real)((d->tau - 0.5) / 3.0);
  if (kind == POL_KIND_INLET) {
    m.rho = (real)d->inlet_rho;
    m.ux = (real)d->inlet_u[0];
    m.uy = (real)d->inlet_u[1];
    m.uz = (real)d->inlet_u[2];
    for (int q = 0; q < 19; ++q) g[q] = feq(q, m.rho, m.ux, m.uy, m.uz);
    return m;
  }
  real rho = 0, jx = 0, jy = 0, jz = 0;
  for (int q = 0; q < 19; ++q) {
    rho += g[q];
    jx += (real)EX[q] * g[q];
    jy += (real)EY[q] * g[q];
    jz += (real)EZ[q] * g[q];
  }
  real inv = rho > (real)1e-12 ? (real)1 / rho : (real)0;
  real F[3] = {F0[0], F0[1], F0[2]};
  real u[3];
  const real B = pd ? pd[3] : (real)0;
  if (porous || B > 0) {
    /* rho u = j + (F0 + BV - B u - rho (nu/K) u - rho (C_F/sqrt K)|u| u)/2:
     * u (c0 + c1 |u|) = v, v = (j + F0/2 + BV/2)/rho, c0 = 1 + B/(2 rho) [+ nu/(2K)],
     * c1 = C_F/(2 sqrt K) [porous]; |u| = 2|v| / (c0 + sqrt(c0^2 + 4 c1 |v|)) */
    real v[3] = {(jx + (real)0.5 * (F[0] + (pd ? pd[0] : 0))) * inv,
                 (jy + (real)0.5 * (F[1] + (pd ? pd[1] : 0))) * inv,
                 (jz + (real)0.5 * (F[2] + (pd ? pd[2] : 0))) * inv};
    real c0 = (real)1 + (real)0.5 * B * inv, c1 = 0, dk = 0, fk = 0;
    if (porous) {
      dk = nu0 / K;
      fk = (real)d->forch_CF / sqrt(K);
      c0 += (real)0.5 * dk;
      c1 = (real)0.5 * fk;
    }
    real vm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    real s = (real)2 / (c0 + sqrt(c0 * c0 + (real)4 * c1 * vm));
    for (int i = 0; i < 3; ++i) u[i] = v[i] * s;
    real un = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    real c = rho * (dk + fk * un) + B;
    for (int i = 0; i < 3; ++i) F[i] += (pd ? pd[i] : 0) - c * u[i];
  } else {
    real v[3] = {(jx + (real)0.5 * F[0]) * inv, (jy + (real)0.5 * F[1]) * inv, (jz + (real)0.5 * F[2]) * inv};
    for (int i = 0; i < 3; ++i) u[i] = v[i];
  }
  m.rho = rho;
  m.ux = u[0];
  m.uy = u[1];
  m.uz = u[2];
  if (kind == POL_KIND_OUTLET) {
    for (int q = 0; q < 19; ++q) g[q] = feq(q, (real)d->outlet_rho, u[0], u[1], u[2]);
    return m;
  }
  real fe[19];
  for (int q = 0; q < 19; ++q) fe[q] = feq(q, rho, u[0], u[1], u[2]);
  real tau = tau0;
  if (les && !porous && nu_fd >= 0) {
    /* given eddy viscosity (the finite-difference form): tau_eff = tau + 3 nu,
     * clamped to [0.55, 1.90] (legacy/lbm_solver.py:565-572) */
    tau = tau0 + (real)3 * nu_fd;
    if (tau < (real)0.55) tau = (real)0.55;
    if (tau > (real)1.90) tau = (real)1.90;
  } else if (les && !porous) {
    real P[3][3] = {{0}};
    for (int q = 0; q < 19; ++q) {
      int e[3] = {EX[q], EY[q], EZ[q]};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) P[a][b] += (real)(e[a] * e[b]) * (g[q] - fe[q]);
    }
    real Q2 = 0;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) Q2 += P[a][b] * P[a][b];
    real Qn = sqrt(Q2);
    real cles = (real)(18.0 * sqrt(2.0) * d->cs_smag * d->cs_smag);
    tau = (real)0.5 * (tau0 + sqrt(tau0 * tau0 + cles * Qn * inv));
    real tmax = (real)(d->tau + 3.0 * d->nu_sgs_max);
    if (tau > tmax) tau = tmax;
  }
  real om = (real)1 / tau;
  real uF = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
  for (int q = 0; q < 19; ++q) {
    real eu = (real)EX[q] * u[0] + (real)EY[q] * u[1] + (real)EZ[q] * u[2];
    real eF = (real)EX[q] * F[0] + (real)EY[q] * F[1] + (real)EZ[q] * F[2];
    real src = ((real)1 - (real)0.5 * om) * wq(q) * ((real)3 * (eF - uF) + (real)9 * eu * eF);
    g[q] = g[q] - om * (g[q] - fe[q]) + src;
  }
  return m;
}

static int wrap_or_clamp(int s, int n, int periodic) {
  if (s >= 0 && s < n) return s;
  if (periodic) return (s + n) % n;
  return s < 0 ? 0 : n - 1; /* clamp onto the cell's own plane */
}

/* Gather the 19 incoming populations of cell (x,y,z). */
static void gather(const pol_desc *d, const uint32_t *flags, const real *fin, int x, int y, int z, real *g) {
  const size_t N = (size_t)d->nx * d->ny * d->nz;
  const size_t cell = ((size_t)z * d->ny + y) * d->nx + x;
  const uint32_t fl = flags ? flags[cell] : 0u;
  for (int q = 0; q < 19; ++q) {
    if (q && (fl & (1u << q))) {
      g[q] = fin[(size_t)opp_of(q) * N + cell];
      continue;
    }
    int sx = wrap_or_clamp(x - EX[q], d->nx, d->options & POL_PERIODIC_X);
    int sy = wrap_or_clamp(y - EY[q], d->ny, d->options & POL_PERIODIC_Y);
    int sz = wrap_or_clamp(z - EZ[q], d->nz, d->options & POL_PERIODIC_Z);
    g[q] = fin[(size_t)q * N + ((size_t)sz * d->ny + sy) * d->nx + sx];
  }
}

/* One step: fin -> fout.  rho_out/u_out (may be NULL) receive the moments of
 * the collided state (solid cells: rho 1, u 0).  red_out (may be NULL)
 * receives the POL_NRED reductions (sums in FP64 in cell order).  pdrag
 * (may be NULL) is the [4][cells] particle-drag field. */
int SFX(oracle_step_nu)(const pol_desc *d, const uint32_t *flags, const real *force, const real *perm,
                        const real *pdrag, const real *nu, const real *fin, real *fout, real *rho_out, real *u_out,
                        double *red_out);
int SFX(oracle_step)(const pol_desc *d, const uint32_t *flags, const real *force, const real *perm, const real *pdrag,
                     const real *fin, real *fout, real *rho_out, real *u_out, double *red_out) {
  return SFX(oracle_step_nu)(d, flags, force, perm, pdrag, NULL, fin, fout, rho_out, u_out, red_out);
}

/* One step with the eddy viscosity given per cell (nu, may be NULL: the
 * local-stress form) when the descriptor has POL_LES and POL_LES_FD. */
int SFX(oracle_step_nu)(const pol_desc *d, const uint32_t *flags, const real *force, const real *perm,
                        const real *pdrag, const real *nu, const real *fin, real *fout, real *rho_out, real *u_out,
                        double *red_out) {
  const size_t N = (size_t)d->nx * d->ny * d->nz;
  const int use_flags = (d->options & POL_USE_FLAGS) && flags;
  double red[POL_NRED];
  for (int s = 0; s < POL_NRED; ++s) red[s] = 0.0;
  red[POL_RED_UMAX] = -1e300;
  red[POL_RED_RHOMAX] = -1e300;
  red[POL_RED_RHOMIN] = 1e300;
  double *zred = (double *)calloc((size_t)d->nz * POL_NRED, sizeof(double));
  if (!zred) return -1;
#pragma omp parallel for schedule(static)
  for (int z = 0; z < d->nz; ++z) {
    double *r = zred + (size_t)z * POL_NRED;
    r[POL_RED_UMAX] = -1e300;
    r[POL_RED_RHOMAX] = -1e300;
    r[POL_RED_RHOMIN] = 1e300;
    for (int y = 0; y < d->ny; ++y)
      for (int x = 0; x < d->nx; ++x) {
        const size_t cell = ((size_t)z * d->ny + y) * d->nx + x;
        const uint32_t fl = use_flags ? flags[cell] : 0u;
        const unsigned kind = (fl >> POL_KIND_SHIFT) & POL_KIND_MASK;
        if (kind == POL_KIND_SOLID) {
          if (rho_out) rho_out[cell] = 1;
          if (u_out) u_out[cell] = u_out[N + cell] = u_out[2 * N + cell] = 0;
          continue;
        }
        real g[19];
        if (kind != POL_KIND_INLET) gather(d, use_flags ? flags : NULL, fin, x, y, z, g);
        real F0[3] = {0, 0, 0};
        if (d->options & POL_GRAVITY)
          for (int i = 0; i < 3; ++i) F0[i] = (real)d->gravity[i];
        if ((d->options & POL_FORCE_FIELD) && force)
          for (int i = 0; i < 3; ++i) F0[i] += force[(size_t)i * N + cell];
        const int porous = (fl & POL_FLAG_POROUS) != 0;
        real K = (real)d->darcy_K;
        if (porous && (d->options & POL_PERM_FIELD) && perm) K = perm[cell];
        const int les = (d->options & POL_LES) && !(fl & POL_FLAG_NO_LES);
        real pd[4];
        const int has_pd = (d->options & POL_PARTICLE_FIELD) && pdrag;
        if (has_pd)
          for (int i = 0; i < 4; ++i) pd[i] = pdrag[(size_t)i * N + cell];
        const real nu_fd = (nu && (d->options & POL_LES_FD)) ? nu[cell] : (real)-1;
        moments m = collide(g, kind, porous, les, K, F0, has_pd ? pd : NULL, d, nu_fd);
        if (fout)
          for (int q = 0; q < 19; ++q) fout[(size_t)q * N + cell] = g[q];
        if (rho_out) rho_out[cell] = m.rho;
        if (u_out) {
          u_out[cell] = m.ux;
          u_out[N + cell] = m.uy;
          u_out[2 * N + cell] = m.uz;
        }
        const double rho = m.rho, ux = m.ux, uy = m.uy, uz = m.uz;
        const double usq = ux * ux + uy * uy + uz * uz;
        const int fin_ok = isfinite(rho) && isfinite(usq);
        r[POL_RED_MASS] += rho;
        r[POL_RED_MOMX] += rho * ux;
        r[POL_RED_MOMY] += rho * uy;
        r[POL_RED_MOMZ] += rho * uz;
        r[POL_RED_KE] += 0.5 * rho * usq;
        if (fin_ok && sqrt(usq) > r[POL_RED_UMAX]) r[POL_RED_UMAX] = sqrt(usq);
        if (!fin_ok) r[POL_RED_NONFINITE] += 1.0;
        if (kind == POL_KIND_OUTLET) r[POL_RED_OUTFLUX] += -rho * uz;
        if (kind == POL_KIND_INLET) r[POL_RED_INFLUX] += -rho * uz;
        r[POL_RED_CELLS] += 1.0;
        if (fin_ok && rho < r[POL_RED_RHOMIN]) r[POL_RED_RHOMIN] = rho;
        if (fin_ok && rho > r[POL_RED_RHOMAX]) r[POL_RED_RHOMAX] = rho;
      }
  }
  for (int z = 0; z < d->nz; ++z) {
    const double *r = zred + (size_t)z * POL_NRED;
    for (int s = 0; s < POL_NRED; ++s) {
      if (s == POL_RED_UMAX || s == POL_RED_RHOMAX) red[s] = r[s] > red[s] ? r[s] : red[s];
      else if (s == POL_RED_RHOMIN) red[s] = r[s] < red[s] ? r[s] : red[s];
      else red[s] += r[s];
    }
  }
  free(zred);
  if (red_out) memcpy(red_out, red, sizeof(red));
  return 0;
}

/* f = feq(rho, u) (rho/u may be NULL -> 1 / 0). */
void SFX(oracle_init_equilibrium)(int nx, int ny, int nz, const real *rho, const real *u, real *f) {
  const size_t N = (size_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)N; ++i) {
    real r = rho ? rho[i] : (real)1;
    real ux = u ? u[i] : 0, uy = u ? u[N + i] : 0, uz = u ? u[2 * N + i] : 0;
    for (int q = 0; q < 19; ++q) f[(size_t)q * N + i] = feq(q, r, ux, uy, uz);
  }
}

/* ---- CPUBackend restatement (cpu_backend.py:79-190) ---------------------- */
static const int CB_EX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
static const int CB_EY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 1, -1, 0, 0, 0, 0, 1, -1};
static const int CB_EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1};
static const int CB_OPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

/* One execute_collision_streaming call.  f (pre-collision, updated in place),
 * f_new (post-collision scratch, persists between calls like the reference's
 * field), rho/u outputs, solid (uint8, may be NULL).  Populations in the
 * CPUBackend's own direction order. */
void SFX(oracle_cpu_backend_step)(int nx, int ny, int nz, real tau, real *f, real *f_new, real *rho, real *u,
                                  const uint8_t *solid) {
  const size_t N = (size_t)nx * ny * nz;
  const real w[19] = {(real)(1.0 / 3.0), (real)(1.0 / 18.0), (real)(1.0 / 18.0), (real)(1.0 / 18.0),
                      (real)(1.0 / 18.0), (real)(1.0 / 18.0), (real)(1.0 / 18.0), (real)(1.0 / 36.0),
                      (real)(1.0 / 36.0), (real)(1.0 / 36.0), (real)(1.0 / 36.0), (real)(1.0 / 36.0),
                      (real)(1.0 / 36.0), (real)(1.0 / 36.0), (real)(1.0 / 36.0), (real)(1.0 / 36.0),
                      (real)(1.0 / 36.0), (real)(1.0 / 36.0), (real)(1.0 / 36.0)};
  /* collision (cpu_backend.py:89-138): index order (i, j, k) = (x, y, z) */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t c = ((size_t)k * ny + j) * nx + i;
        if (solid && solid[c]) continue;
        real r = 0, ux = 0, uy = 0, uz = 0;
        for (int q = 0; q < 19; ++q) {
          real fq = f[(size_t)q * N + c];
          r += fq;
          ux += fq * (real)CB_EX[q];
          uy += fq * (real)CB_EY[q];
          uz += fq * (real)CB_EZ[q];
        }
        if (r > (real)1e-12) {
          real ir = (real)1 / r;
          ux *= ir;
          uy *= ir;
          uz *= ir;
        } else {
          ux = uy = uz = 0;
        }
        rho[c] = r;
        u[c] = ux;
        u[N + c] = uy;
        u[2 * N + c] = uz;
        real usq = ux * ux + uy * uy + uz * uz;
        for (int q = 0; q < 19; ++q) {
          real eu = (real)CB_EX[q] * ux + (real)CB_EY[q] * uy + (real)CB_EZ[q] * uz;
          real fe = w[q] * r * ((real)1 + (real)3 * eu + (real)4.5 * eu * eu - (real)1.5 * usq);
          real fo = f[(size_t)q * N + c];
          real fn = fo - (fo - fe) / tau;
          if (fn < 0) fn = 0;
          f_new[(size_t)q * N + c] = fn;
        }
      }
  /* streaming (cpu_backend.py:147-166): pull, out-of-box source keeps own f */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t c = ((size_t)k * ny + j) * nx + i;
        for (int q = 0; q < 19; ++q) {
          int si = i - CB_EX[q], sj = j - CB_EY[q], sk = k - CB_EZ[q];
          if (si >= 0 && si < nx && sj >= 0 && sj < ny && sk >= 0 && sk < nz)
            f[(size_t)q * N + c] = f_new[(size_t)q * N + ((size_t)sk * ny + sj) * nx + si];
        }
      }
  /* solid pass (cpu_backend.py:172-183): sequential f[opp[q]] = f[q] */
  if (solid) {
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < (long long)N; ++c) {
      if (!solid[c]) continue;
      for (int q = 0; q < 19; ++q) f[(size_t)CB_OPP[q] * N + c] = f[(size_t)q * N + c];
    }
  }
}

/* ---- particles (coffee_particles.py) ------------------------------------- */
/* i = (int)max(0, min(n-2, p)) in float32, frac = clamp(p - i, 0, 1) */
void SFX(oracle_particle_index)(const float *pos, int n, int nx, int ny, int nz, int32_t *idx, float *frac) {
  const int dims[3] = {nx, ny, nz};
  for (int d = 0; d < 3; ++d)
    for (int p = 0; p < n; ++p) {
      volatile float pp = pos[(size_t)d * n + p];
      float hi = (float)(dims[d] - 2);
      float c = fmaxf(0.0f, fminf(hi, pp));
      int i = (int)c;
      float t = pp - (float)i;
      idx[(size_t)d * n + p] = i;
      frac[(size_t)d * n + p] = fmaxf(0.0f, fminf(1.0f, t)) + 0.0f; /* -0 -> +0 */
    }
}

static real drag_cd(real re) {
  real cd = (real)24 / (re > (real)0.01 ? re : (real)0.01);
  if (re >= (real)0.1 && re < (real)1000) cd = (real)24 / re * ((real)1 + (real)0.15 * pow(re, (real)0.687));
  else if (re >= (real)1000) cd = (real)0.44;
  return cd;
}
real SFX(oracle_drag_coefficient)(real re) { return drag_cd(re); }

typedef struct {
  int count;
  int drag_law; /* POL_DRAG_TEXTBOOK: |F| ~ |u_rel|^2; POL_DRAG_REFERENCE: ~ |u_rel| as coded in the reference */
  real rho_fluid, nu_fluid, drag_max_per_mass, alpha, g[3], dt, max_disp, wall_damping;
} oracle_particle_params;

/* Coupling pass of one particle set (coffee_particles.py:1107-1154).
 * Gather u at the particle, drag = beta (u_f - v) with the Schiller-Naumann
 * C_D, then the reaction:
 *   explicit = 0: accumulate the [4][cells] drag field (w beta v, w beta) in
 *                 `field` (the build's implicit form);
 *   explicit = 1: accumulate w (-drag_new) in the [3][cells] reaction field
 *                 `field`, as distribute_force_to_grid does
 *                 (coffee_particles.py:1146-1149,1049-1086).
 * relax = 1 also applies the under-relaxation (coffee_particles.py:1200-1212). */
static void couple_pass(int nx, int ny, int nz, const oracle_particle_params *pp, const real *pos, const real *vel,
                        const real *radius, const real *mass, const int32_t *active, const real *u_field, real *field,
                        real *u_fluid, real *drag_new, real *drag_old, real *drag, real *re_out, real *cd_out,
                        int explicit_reaction, int relax) {
  const size_t N = (size_t)nx * ny * nz;
  const int P = pp->count;
  const size_t nxy = (size_t)nx * ny;
  for (int p = 0; p < P; ++p) {
    if (!active[p]) continue;
    int ii[3];
    real fr[3];
    const int dims[3] = {nx, ny, nz};
    for (int d = 0; d < 3; ++d) {
      /* index formula of coffee_particles.py:1159-1174 in the state's type */
      real pr = pos[(size_t)d * P + p];
      real hi = (real)(dims[d] - 2);
      real c = pr < hi ? pr : hi;
      c = c > 0 ? c : 0;
      if (pr != pr) c = hi; /* NaN: min returns the number operand */
      ii[d] = (int)c;
      real t = pr - (real)ii[d];
      fr[d] = t < 0 ? 0 : (t > 1 ? 1 : t);
    }
    const size_t c000 = (size_t)ii[2] * nxy + (size_t)ii[1] * nx + ii[0];
    const size_t cs[8] = {c000, c000 + nxy, c000 + nx, c000 + nx + nxy, c000 + 1, c000 + nxy + 1, c000 + nx + 1,
                          c000 + nx + nxy + 1};
    const real fx = fr[0], fy = fr[1], fz = fr[2];
    const real ws[8] = {(1 - fx) * (1 - fy) * (1 - fz), (1 - fx) * (1 - fy) * fz, (1 - fx) * fy * (1 - fz),
                        (1 - fx) * fy * fz, fx * (1 - fy) * (1 - fz), fx * (1 - fy) * fz, fx * fy * (1 - fz),
                        fx * fy * fz};
    real uf[3];
    for (int d = 0; d < 3; ++d) {
      real s = 0;
      for (int k = 0; k < 8; ++k) s += ws[k] * u_field[(size_t)d * N + cs[k]];
      uf[d] = s;
      u_fluid[(size_t)d * P + p] = s;
    }
    real rel[3] = {uf[0] - vel[p], uf[1] - vel[P + p], uf[2] - vel[2 * P + p]};
    real mag = sqrt(rel[0] * rel[0] + rel[1] * rel[1] + rel[2] * rel[2]);
    real dn[3] = {0, 0, 0};
    if (mag > (real)1e-8) {
      real r = radius[p];
      real mu = pp->rho_fluid * pp->nu_fluid;
      real re = pp->rho_fluid * mag * 2 * r / (mu > (real)1e-8 ? mu : (real)1e-8);
      real cd = drag_cd(re);
      re_out[p] = re;
      cd_out[p] = cd;
      real area = (real)3.14159 * r * r;
      real dm = (real)0.5 * pp->rho_fluid * cd * area * mag;
      if (pp->drag_law == POL_DRAG_TEXTBOOK) dm *= mag;
      real cap = mass[p] * pp->drag_max_per_mass;
      if (dm > cap) dm = cap;
      const real beta = dm / mag;
      for (int d = 0; d < 3; ++d) dn[d] = beta * rel[d];
      if (explicit_reaction) {
        for (int d = 0; d < 3; ++d)
          for (int k = 0; k < 8; ++k) field[(size_t)d * N + cs[k]] += ws[k] * (-dn[d]);
      } else {
        /* drag field: w beta v (3 components) and w beta */
        const real comp[4] = {beta * vel[p], beta * vel[P + p], beta * vel[2 * P + p], beta};
        for (int d = 0; d < 4; ++d)
          for (int k = 0; k < 8; ++k) field[(size_t)d * N + cs[k]] += ws[k] * comp[d];
      }
    } else {
      re_out[p] = 0;
      cd_out[p] = 0;
    }
    for (int d = 0; d < 3; ++d) {
      drag_new[(size_t)d * P + p] = dn[d];
      if (relax) {
        real rl = pp->alpha * dn[d] + (1 - pp->alpha) * drag_old[(size_t)d * P + p];
        drag[(size_t)d * P + p] = rl;
        drag_old[(size_t)d * P + p] = rl;
      }
    }
  }
}

/* couple: the implicit drag field (w beta v, w beta) in `force`, relaxed. */
void SFX(oracle_particles_couple)(int nx, int ny, int nz, const oracle_particle_params *pp, const real *pos,
                                  const real *vel, const real *radius, const real *mass, const int32_t *active,
                                  const real *u_field, real *force, real *u_fluid, real *drag_new, real *drag_old,
                                  real *drag, real *re_out, real *cd_out) {
  couple_pass(nx, ny, nz, pp, pos, vel, radius, mass, active, u_field, force, u_fluid, drag_new, drag_old, drag,
              re_out, cd_out, 0, 1);
}

/* compute_two_way_coupling_forces (coffee_particles.py:1107-1154) without
 * the relaxation: explicit_reaction selects the reference's reaction field
 * ([3][cells], accumulated: the caller clears it as :1109-1111 does) or the
 * implicit drag field ([4][cells]). */
void SFX(oracle_particles_compute)(int nx, int ny, int nz, const oracle_particle_params *pp, const real *pos,
                                   const real *vel, const real *radius, const real *mass, const int32_t *active,
                                   const real *u_field, real *field, real *u_fluid, real *drag_new, real *re_out,
                                   real *cd_out, int explicit_reaction) {
  couple_pass(nx, ny, nz, pp, pos, vel, radius, mass, active, u_field, field, u_fluid, drag_new, NULL, NULL, re_out,
              cd_out, explicit_reaction, 0);
}

/* apply_under_relaxation (coffee_particles.py:1200-1212), active particles. */
void SFX(oracle_particles_relax)(int count, real alpha, const int32_t *active, const real *drag_new, real *drag_old,
                                 real *drag) {
  for (int p = 0; p < count; ++p) {
    if (!active[p]) continue;
    for (int d = 0; d < 3; ++d) {
      const real rl = alpha * drag_new[(size_t)d * count + p] + (1 - alpha) * drag_old[(size_t)d * count + p];
      drag[(size_t)d * count + p] = rl;
      drag_old[(size_t)d * count + p] = rl;
    }
  }
}

/* advance: clear the footprint, implicit linearised drag, clamp, walls. */
void SFX(oracle_particles_advance)(int nx, int ny, int nz, const oracle_particle_params *pp, const uint32_t *flags,
                                   real *pos, real *vel, const real *mass, int32_t *active, const real *u_fluid,
                                   const real *drag_new, real *force) {
  const size_t N = (size_t)nx * ny * nz;
  const size_t nxy = (size_t)nx * ny;
  const int P = pp->count;
  for (int p = 0; p < P; ++p) {
    if (!active[p]) continue;
    const int dims[3] = {nx, ny, nz};
    int ii[3];
    for (int d = 0; d < 3; ++d) {
      real pr = pos[(size_t)d * P + p];
      real hi = (real)(dims[d] - 2);
      real c = pr < hi ? pr : hi;
      c = c > 0 ? c : 0;
      if (pr != pr) c = hi;
      ii[d] = (int)c;
    }
    const size_t c000 = (size_t)ii[2] * nxy + (size_t)ii[1] * nx + ii[0];
    const size_t cs[8] = {c000, c000 + 1, c000 + nx, c000 + nx + 1, c000 + nxy, c000 + nxy + 1, c000 + nxy + nx,
                          c000 + nxy + nx + 1};
    if (force)
      for (int d = 0; d < 4; ++d)
        for (int k = 0; k < 8; ++k) force[(size_t)d * N + cs[k]] = 0;
    real v[3], uf[3], dn[3], x[3];
    for (int d = 0; d < 3; ++d) {
      v[d] = vel[(size_t)d * P + p];
      uf[d] = u_fluid[(size_t)d * P + p];
      dn[d] = drag_new[(size_t)d * P + p];
      x[d] = pos[(size_t)d * P + p];
    }
    real rel[3] = {uf[0] - v[0], uf[1] - v[1], uf[2] - v[2]};
    real mag = sqrt(rel[0] * rel[0] + rel[1] * rel[1] + rel[2] * rel[2]);
    real dmag = sqrt(dn[0] * dn[0] + dn[1] * dn[1] + dn[2] * dn[2]);
    real beta = mag > (real)1e-8 ? dmag / mag : 0;
    real h = pp->dt * beta / mass[p];
    for (int d = 0; d < 3; ++d) v[d] = (v[d] + h * uf[d] + pp->dt * pp->g[d]) / (1 + h);
    real dx[3] = {pp->dt * v[0], pp->dt * v[1], pp->dt * v[2]};
    real dl = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
    if (dl > pp->max_disp) {
      real s = pp->max_disp / dl;
      for (int d = 0; d < 3; ++d) dx[d] *= s;
    }
    real xn[3] = {x[0] + dx[0], x[1] + dx[1], x[2] + dx[2]};
    if (!(xn[2] >= 1) || !(xn[0] >= 0) || !(xn[1] >= 0) || !(xn[0] < nx - 1) || !(xn[1] < ny - 1) ||
        !(xn[2] < nz - 1)) {
      active[p] = 0;
    } else {
      size_t c = (size_t)(int)xn[2] * nxy + (size_t)(int)xn[1] * nx + (size_t)(int)xn[0];
      uint32_t fl = flags ? flags[c] : 0u;
      unsigned kd = (fl >> POL_KIND_SHIFT) & POL_KIND_MASK;
      if (kd == POL_KIND_OUTLET) {
        active[p] = 0;
      } else if (kd == POL_KIND_SOLID || (fl & POL_FLAG_POROUS)) {
        for (int d = 0; d < 3; ++d) {
          xn[d] = x[d];
          v[d] *= pp->wall_damping;
        }
      }
    }
    for (int d = 0; d < 3; ++d) {
      pos[(size_t)d * P + p] = xn[d];
      vel[(size_t)d * P + p] = v[d];
    }
  }
}
