// The fused D3Q19 stream-collide kernel template, shared by the translation
// units that instantiate its dense-layout (lbm_step.hip) and compact-layout
// (lbm_step_sparse.hip) variants.  See lbm_step.hip for the algorithm.
#pragma once
#include "cell_model.h"
#include "compact_gather.h"
#include "engine.h"

namespace pol {

template <typename R>
__device__ inline Phys<R> make_phys(const StepArgs& a) {
  Phys<R> P;
  if constexpr (sizeof(R) == 4) {
    const PhysF& p = a.pf;  // rounded on the host: no per-lane conversions
    P.tau0 = p.tau0;
    P.omega0 = p.omega0;
    P.nu0 = p.nu0;
    P.les_c = p.les_c;
    P.tau_max = p.tau_max;
    P.G[0] = p.G[0];
    P.G[1] = p.G[1];
    P.G[2] = p.G[2];
    P.K = p.K;
    P.CF = p.CF;
    P.inlet_rho = p.inlet_rho;
    P.inlet_u[0] = p.inlet_u[0];
    P.inlet_u[1] = p.inlet_u[1];
    P.inlet_u[2] = p.inlet_u[2];
    P.outlet_rho = p.outlet_rho;
  } else {
    P.tau0 = (R)a.tau0;
    P.omega0 = R(1) / (R)a.tau0;
    P.nu0 = (R)a.nu0;
    P.les_c = (R)a.les_c;
    P.tau_max = (R)a.tau_max;
    P.G[0] = (R)a.G[0];
    P.G[1] = (R)a.G[1];
    P.G[2] = (R)a.G[2];
    P.K = (R)a.K;
    P.CF = (R)a.CF;
    P.inlet_rho = (R)a.inlet_rho;
    P.inlet_u[0] = (R)a.inlet_u[0];
    P.inlet_u[1] = (R)a.inlet_u[1];
    P.inlet_u[2] = (R)a.inlet_u[2];
    P.outlet_rho = (R)a.outlet_rho;
  }
  return P;
}

// COLLIDE = false: collisionless streaming (f_out = the gathered populations,
// bounce-back links and open faces as in a step), the step's memory access
// pattern with none of its arithmetic.
template <typename S, bool FLAGS, bool FULL, bool STORE, bool A32, bool NT = false, bool SPARSE = false, bool XF = true,
          bool COLLIDE = true>
__global__ void __launch_bounds__(BLOCK) k_stream_collide(const StepArgs a) {
  static_assert(!SPARSE || (FLAGS && FULL), "the compact layout runs the general kernel");
  using C = Codec<S>;
  using R = typename C::real;
  using AD = Addr<A32>;
  using AI = typename AD::idx;
  const uint32_t lblk = ordered_block(a.order, a.per_xcd);
  if (lblk >= a.nblk) return;  // whole padding block exits (uniform)
  uint32_t cell, x, y, z;
  int rb[9];          // compact layout: slot bases of the nine neighbour rows
  int sx = 0;         // compact layout: the lane's unwrapped x (its slot is rb[4] + sx)
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

  const uint32_t fl = FLAGS ? (inside ? (plain ? 0u : a.flags[cell]) : (POL_KIND_SOLID << POL_KIND_SHIFT)) : 0u;
  const uint32_t kind = (fl >> POL_KIND_SHIFT) & POL_KIND_MASK;
  const bool active = inside && kind != POL_KIND_SOLID;

  const Phys<R> P = make_phys<R>(a);

  const S* __restrict__ fin = (const S*)a.fin;
  S* __restrict__ fout = (S*)a.fout;
  const uint32_t N = a.N;  // dense cells (flags, fields, rho/u)
  const uint32_t PS = a.P;  // population plane stride
  // population slot of this cell: the dense index, or its compact-layout slot
  uint32_t own = cell;
  if constexpr (SPARSE) own = slot_clamp(rb[4] + sx, PS);

  R f[Q];  // scaled deviations 36 (f - w)
  // FP16 compact launches keep the loaded words and let the collision widen
  // them inside its first operations (collide_cell_lp<.., true>)
  constexpr bool MIX = SPARSE && COLLIDE && std::is_same_v<S, _Float16>;
  uint32_t h[Q];
  CellOut<R> o{R(0), R(0), R(0), R(0)};
  if (active) {
    if constexpr (!COLLIDE) {
#pragma unroll
      for (int q = 0; q < Q; ++q) f[q] = R(0);  // inlet lanes gather nothing
    }
    if constexpr (SPARSE) {
      if (kind != POL_KIND_INLET) {
        if constexpr (MIX) {
          compact_sources<S, A32, XF, true>(fin, PS, rb, x, a.nx, (a.options & POL_PERIODIC_X) != 0, own, fl,
                                            [&](auto qc, S v) {
                                              h[decltype(qc)::value] = __builtin_bit_cast(uint16_t, v);
                                            });
          f[0] = (R)__builtin_bit_cast(_Float16, (uint16_t)h[0]);
        } else {
          compact_sources<S, A32, XF, false, COLLIDE && sizeof(S) == 2>(
              fin, PS, rb, x, a.nx, (a.options & POL_PERIODIC_X) != 0, own, fl,
              [&](auto qc, R v) { f[decltype(qc)::value] = v; });
        }
      }
    } else if (!FULL || kind != POL_KIND_INLET) {
      // neighbour offsets; non-periodic axes clamp to the cell's own plane
      // (zero-gradient for open-boundary cells; walls use bounce-back bits)
      const bool px = a.options & POL_PERIODIC_X, py = a.options & POL_PERIODIC_Y, pz = a.options & POL_PERIODIC_Z;
      const int nx = (int)a.nx, nxy = (int)a.nxy;
      const int oxm = x == 0 ? (px ? nx - 1 : 0) : -1;               // source at x-1
      const int oxp = x == a.nx - 1 ? (px ? -(nx - 1) : 0) : 1;      // source at x+1
      const int oym = y == 0 ? (py ? ((int)a.ny - 1) * nx : 0) : -nx;
      const int oyp = y == a.ny - 1 ? (py ? -((int)a.ny - 1) * nx : 0) : nx;
      const int ozm = z == 0 ? (pz ? ((int)a.nz - 1) * nxy : 0) : -nxy;
      const int ozp = z == a.nz - 1 ? (pz ? -((int)a.nz - 1) * nxy : 0) : nxy;
      static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int off = (EX(q) > 0 ? oxm : (EX(q) < 0 ? oxp : 0)) + (EY(q) > 0 ? oym : (EY(q) < 0 ? oyp : 0)) +
                        (EZ(q) > 0 ? ozm : (EZ(q) < 0 ? ozp : 0));
        AI src = (AI)q * PS + (uint32_t)((int)cell + off);
        if constexpr (FLAGS && q != 0) {
          if (fl & (1u << q)) src = (AI)OPP(q) * PS + own;
        }
        // bounce-back reads D*_{opp q}(x); w_{opp q} = w_q, so the stored
        // deviation of direction opp(q) is the deviation of link q.
        f[q] = C::load(AD::at(fin, src));
      });
    }
    R fx = R(0), fy = R(0), fz = R(0), K = P.K;
    R bx = R(0), by = R(0), bz = R(0), B = R(0);
    if constexpr (FULL) {
      if (a.options & POL_FORCE_FIELD) {
        const R* F = (const R*)a.force;
        fx = F[cell];
        fy = F[(size_t)N + cell];
        fz = F[2 * (size_t)N + cell];
      }
      if (a.options & POL_PARTICLE_FIELD) {
        const R* Pd = (const R*)a.pdrag;
        bx = Pd[cell];
        by = Pd[(size_t)N + cell];
        bz = Pd[2 * (size_t)N + cell];
        B = Pd[3 * (size_t)N + cell];
      }
    }
    const bool porous = FULL && (fl & POL_FLAG_POROUS);
    if constexpr (FULL) {
      if (porous && (a.options & POL_PERM_FIELD)) K = ((const R*)a.perm)[cell];
    }
    const bool les = FULL && (a.options & POL_LES) && !(fl & POL_FLAG_NO_LES);
    // moments-only passes use the same velocity definition the collision uses
    const bool les_on = STORE && les;
    // the reference's finite-difference eddy viscosity, formed beforehand
    R nu_fd = R(-1);
    if constexpr (FULL && STORE) {
      if (les_on && (a.options & POL_LES_FD)) nu_fd = ((const R*)a.nu_sgs)[cell];
    }
    if constexpr (!COLLIDE)
      ;
    else if constexpr (sizeof(S) == 2)
      o = collide_cell_lp<FULL, MIX>(f, kind, porous, les_on, K, fx, fy, fz, bx, by, bz, B, P, h, nu_fd);
    else
      o = collide_cell<R, FULL>(f, kind, porous, les_on, K, fx, fy, fz, bx, by, bz, B, P, nu_fd);
  }
  if constexpr (STORE) {
    // Solid lanes next to fluid ones store zero deviations (f = w; no gather
    // ever reads a solid cell), so that every 128-byte line of the output is
    // written whole: a partly written line costs HBM a read-modify-write.
    // Chunk launches list only chunks holding fluid, so all their lanes store;
    // dense launches store a solid lane when its 32-lane group holds fluid;
    // the compact layout stores exactly the slots of the stored runs.
    bool store = active;
    if constexpr (FLAGS) {
      if constexpr (SPARSE) {
        store = in_run;
      } else {
        const uint64_t any = __ballot(active);
        const uint64_t group = 0xffffffffull << (threadIdx.x & 32);
        store = inside && (a.chunks != nullptr || (any & group) != 0);
      }
      if (!active) {
#pragma unroll
        for (int q = 0; q < Q; ++q) f[q] = R(0);
      }
    }
    if (store) {
      // per-plane scalar base + the lane's one slot offset
      if constexpr (std::is_same_v<S, _Float16>) {
        // FP16 storage: planes in pairs, one packed conversion per pair
        static_for<0, Q / 2>([&](auto pc) {
          constexpr int q = 2 * decltype(pc)::value;
          C::template store2<NT>(AD::at(fout + (AI)q * PS, own), AD::at(fout + (AI)(q + 1) * PS, own), f[q], f[q + 1]);
        });
        if constexpr (NT) C::store_nt(AD::at(fout + (AI)(Q - 1) * PS, own), f[Q - 1]);
        else C::store(AD::at(fout + (AI)(Q - 1) * PS, own), f[Q - 1]);
      } else {
        static_for<0, Q>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          if constexpr (NT) C::store_nt(AD::at(fout + (AI)q * PS, own), f[q]);
          else C::store(AD::at(fout + (AI)q * PS, own), f[q]);
        });
      }
    }
  }

  if (a.what & POL_WANT_MACRO) {
    if (inside) {
      R* rho = (R*)a.rho;
      R* u = (R*)a.u;
      rho[cell] = R(1) + o.rd;
      u[cell] = o.ux;
      u[(size_t)N + cell] = o.uy;
      u[2 * (size_t)N + cell] = o.uz;
    }
  }
  if (a.what & POL_WANT_REDUCE) {
    double v[POL_NRED];
#pragma unroll
    for (int s = 0; s < POL_NRED; ++s) v[s] = red_identity(s);
    if (active) {
      const double rho = 1.0 + (double)o.rd, ux = (double)o.ux, uy = (double)o.uy, uz = (double)o.uz;
      const double usq = ux * ux + uy * uy + uz * uz;
      const bool finite = isfinite(rho) && isfinite(usq);
      v[POL_RED_MASS] = rho;
      v[POL_RED_MOMX] = rho * ux;
      v[POL_RED_MOMY] = rho * uy;
      v[POL_RED_MOMZ] = rho * uz;
      v[POL_RED_KE] = 0.5 * rho * usq;
      v[POL_RED_UMAX] = finite ? sqrt(usq) : -1.0e300;
      v[POL_RED_NONFINITE] = finite ? 0.0 : 1.0;
      v[POL_RED_OUTFLUX] = kind == POL_KIND_OUTLET ? -rho * uz : 0.0;
      v[POL_RED_CELLS] = 1.0;
      v[POL_RED_RHOMIN] = finite ? rho : 1.0e300;
      v[POL_RED_RHOMAX] = finite ? rho : -1.0e300;
      v[POL_RED_INFLUX] = kind == POL_KIND_INLET ? -rho * uz : 0.0;
    }
    block_reduce_store<BLOCK / 64>(v, a.partial, lblk);
  }
}

// compact-layout launches (lbm_step_sparse.hip)
hipError_t launch_stream_collide_sparse(int storage, const StepArgs& a, bool store_f, const Grid& g, hipStream_t s);
hipError_t launch_stream_only_sparse(int storage, const StepArgs& a, const Grid& g, hipStream_t s);

}  // namespace pol
