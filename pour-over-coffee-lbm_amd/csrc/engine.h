// Internal host-side interface between the C ABI (abi.hip) and the kernel
// translation units.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include "lbm_common.h"

namespace pol {

// Everything one stream-collide launch needs; physics kept in FP64 and
// narrowed to the arithmetic type inside the kernel.
// The physics parameters rounded once to FP32 on the host (the FP32 kernels
// read these instead of converting the doubles in every lane).
struct PhysF {
  float tau0, omega0, nu0, les_c, tau_max;
  float G[3];
  float K, CF;
  float inlet_rho, inlet_u[3], outlet_rho;
};

struct StepArgs {
  const void* fin;
  void* fout;
  const uint32_t* flags;
  const void* force;  // real [3][N]
  const void* perm;   // real [N]
  const void* pdrag;  // real [4][N] particle-drag field (sum w beta v, sum w beta)
  const void* nu_sgs; // real [N] finite-difference eddy viscosity (POL_LES_FD), or null
  void* rho;          // real [N]
  void* u;            // real [3][N]
  double* partial;    // [nblk][POL_NRED]
  const uint32_t* chunks;  // chunk / piece list: item -> chunks[item >> cshift] + lane (null: item = cell)
  const uint32_t* records; // compact layout: one row record (POL_RECORD_WORDS) per dense row, else null
  uint32_t P;              // population plane stride: N (dense) or the compact layout's stored cells
  uint32_t n_items;        // number of work items (cells, or chunk cells x chunks)
  uint32_t cshift;         // log2 of the cells per chunk
  uint32_t stream_store;   // select the non-temporal-store kernels (output far above the last-level cache)
  uint32_t xface;          // compact layout: some run reaches x = 0 or nx - 1 (gather wraps / clamps x there)
  uint32_t N, nx, ny, nz, nxy;
  FastDiv dnx, dny;
  uint32_t per_xcd, nblk;
  uint32_t order;          // block order of the stream-collide launches (ordered_block)
  uint32_t options;
  uint32_t what;
  double tau0, nu0, les_c, tau_max;
  double G[3];
  double K, CF;
  double inlet_rho, inlet_u[3], outlet_rho;
  PhysF pf;
};

// Finite-difference eddy viscosity of the reference (les_fd.hip).
struct LesFdArgs {
  const void* u;          // real [3][N]
  const uint32_t* flags;  // cell flags or null (LES everywhere)
  const void* phase;      // real [N] or null (phi = 1)
  void* nu;               // real [N] out
  uint32_t nx, ny, nz, nxy, N;
  FastDiv dnx, dny;
  float cs2;              // (C_s * 1)^2, rounded once from double
  float nu_max;
};
hipError_t launch_les_fd(bool dbl, const LesFdArgs& a, hipStream_t s);

// Launch one pass.  store_f: write post-collision populations (a step);
// false = moments only (macro/reduce of the current state).
hipError_t launch_stream_collide(int storage, const StepArgs& a, bool store_f, const Grid& g, hipStream_t s);
// Collisionless streaming pass a.fin -> a.fout (the step's gather, bounce-back
// links, open faces and stores, no collision, no moments).
hipError_t launch_stream_only(int storage, const StepArgs& a, const Grid& g, hipStream_t s);
// Fold per-block partials into out[POL_NRED] (device), deterministic order.
hipError_t launch_reduce_final(const double* partial, uint32_t nblk, double* out, hipStream_t s);
// f = feq(rho, u) (rho/u real-typed, may be null); a.records selects the
// compact layout (a.fout = f, a.N, a.P, a.n_items, a.cshift, a.dnx, a.nx used)
hipError_t launch_init_equilibrium(int storage, const StepArgs& a, const void* rho, const void* u, const Grid& g,
                                   hipStream_t s);

// Field statistics (diagnostics.hip): one pass over rho / u (+ flags, the
// previous rho / u and a phase field when given) on a (tiles, nz) grid.
struct StatsArgs {
  const void* rho;
  const void* u;
  const uint32_t* flags;  // may be null (no solids)
  const void* prev_rho;   // may be null
  const void* prev_u;     // may be null
  const void* phase;      // may be null (every cell water)
  uint32_t nx, ny, nz, nxy, N;
  FastDiv dnx;
  uint32_t tiles;         // blocks per z plane
  int32_t z_top, z_mid, z_out;
  double* partial;        // [nz * tiles][POL_NSTAT]
  double* zpart;          // [nz * tiles]
};
// Statistics into out_dev[POL_NSTAT] and the per-plane density sums into
// zout_dev[nz] (device), deterministic block order.
hipError_t launch_field_stats(bool real_double, const StatsArgs& s, double* out_dev, double* zout_dev, hipStream_t st);
// NaN / Inf populations of a.fin over the non-solid cells of the launch list
// into out_dev[POL_STAT_FLUID_POP_NAN], [POL_STAT_FLUID_POP_INF].
hipError_t launch_pop_nonfinite(int storage, const StepArgs& a, double* partial, double* out_dev, const Grid& g,
                                hipStream_t st);
// Emergency stabilisation of rho / u and of the populations in a.fout.
hipError_t launch_stabilize(int storage, const StepArgs& a, float max_u, float min_rho, float max_rho, const Grid& g,
                            hipStream_t st);
// |u| and pressure = rho / 3 fields (real [N]).
hipError_t launch_export_fields(bool real_double, const void* rho, const void* u, uint32_t N, void* mag, void* pres,
                                hipStream_t st);

// Phase-field producer (multiphase.hip), dense over every cell.
struct MultiphaseArgs {
  const float* phi;      // current phase field
  float* phi_out;        // next phase field (step)
  const float* phi_new;  // face values the step copies
  float* mu;
  const void* u;         // real [3][N]
  void* rho;             // real [N]
  void* phase;           // real [N]
  void* force;           // real [3][N] body force (apply)
  const uint32_t* flags; // cell flags or null (no solids)
  float* grad_phi;       // optional outputs (interface kernel)
  float* grad_mu;
  float* normal;
  float* curvature;
  float* sforce;
  uint32_t nx, ny, nz, nxy, N;
  FastDiv dnx, dny;
  float sigma, mobility, dt, kappa, rho_air, drho;
  uint32_t apply;
};
hipError_t launch_mp_interface(bool dbl, const MultiphaseArgs& a, bool out, hipStream_t s);
hipError_t launch_mp_phase(bool dbl, const MultiphaseArgs& a, hipStream_t s);
hipError_t launch_mp_density(bool dbl, const MultiphaseArgs& a, hipStream_t s);
hipError_t launch_mp_chemical_potential(const MultiphaseArgs& a, float* lap_out, hipStream_t s);

// Body-force producers (forcing.hip).
struct ForcingArgs {
  void* force;            // real [3][N] body force (or null)
  void* rho;              // real [N]
  const uint32_t* flags;  // cell flags or null (no solids)
  uint32_t nx, ny, nz, nxy, N;
  FastDiv dnx, dny;
};
struct PourArgs {
  float cx, cy, radius, zmin, velocity, flow_rate, dt;
  int32_t height;
  int32_t x0, y0, z0, bx, by, bz;  // launched box
};
struct FilterParticleArgs {
  const uint8_t* zone;
  float* acc;
  const float* pos;  // [3][count]
  float* vel;        // [3][count]
  const int32_t* active;
  uint32_t count, nx, ny, nz, seed;
  float scale;
};
hipError_t launch_pour(bool dbl, const ForcingArgs& f, const PourArgs& p, float* phi, hipStream_t s);
hipError_t launch_density_drive(bool dbl, const ForcingArgs& f, const float* target, float rate, bool half,
                                float max_adj, float lo, float hi, hipStream_t s);
// st: the bound phase field whose pre-collision surface tension follows in
// the same pass (modes 1 and 2), or null
hipError_t launch_pressure_force(bool dbl, const ForcingArgs& f, const MultiphaseArgs* st, float* pf, float max_force,
                                 float cs2, int mode, hipStream_t s);
hipError_t launch_filter_blockage(const uint32_t* cells, uint32_t ncells, float* acc, float* blockage, hipStream_t s);
hipError_t launch_filter_block(const FilterParticleArgs& a, hipStream_t s);

// Thermal D3Q7 state (thermal.hip), dense over every cell.
struct ThermalArgs {
  const void* gin;
  void* gout;
  const void* T;     // temperature of the state being stepped
  void* Tout;        // temperature after the step
  void* hs;          // heat source
  const void* hsb;   // base heat source read instead of hs (reset fused), or null
  const void* u;     // [3][N] or null
  const void* k;     // conductivity or null
  uint32_t nx, ny, nz, nxy, N;
  FastDiv dnx, dny;
  float omega_f;
  double omega_d;
  uint32_t conv;
  double* partial;   // per-block partials (stability counts, statistics)
};
// nsteps steps ping-ponging g[gcur] / T[tcur]; base supplies the fields.
hipError_t launch_thermal_steps(bool dbl, const ThermalArgs& base, void* const g[2], void* const T[2], int& gcur,
                                int& tcur, int nsteps, hipStream_t s);
// Fold per-block partials: nval 1 = sum into out[0]; 3 = {min, max, sum}.
hipError_t launch_thermal_fold(const double* partial, uint32_t nblk, int nval, double* out, hipStream_t s);
hipError_t launch_thermal_stats(bool dbl, const void* T, uint32_t N, double* partial, hipStream_t s);
hipError_t launch_thermal_flux(bool dbl, const ThermalArgs& a, void* flux, hipStream_t s);
hipError_t launch_thermal_init(bool dbl, const ThermalArgs& a, hipStream_t s);
hipError_t launch_thermal_boundary(bool dbl, const ThermalArgs& a, int kind, const uint8_t* mask, double va,
                                   double vb, hipStream_t s);

// Particle state in the real type (float, or double with real_is_double).
struct ParticleArgs {
  int32_t count;
  int32_t real_is_double;
  int32_t drag_law;
  uint32_t nx, ny, nz, nxy, N;
  const uint32_t* flags;
  void *pos, *vel, *radius, *mass;
  int32_t* active;
  void *drag_new, *drag_old, *drag, *u_fluid, *re_p, *cd;
  void* drag_field;  // [4][N] implicit drag field, or null (explicit reaction)
  void* reaction;    // [3][N] explicit reaction field, or null (implicit drag field)
  const void* u_field;
  double rho_fluid, nu_fluid, drag_max_per_mass, alpha;
  double g[3];
  double dt, max_disp, wall_damping;
};
// Coupling pass.  relax: under-relax the drag in the same pass (alpha of the
// arguments); otherwise drag / drag_old are left alone.  With a.reaction the
// pass scatters -drag into it (the field is not cleared here), else it
// scatters (w beta v, w beta) into the drag field.
hipError_t launch_particles_couple(const ParticleArgs& a, bool relax, hipStream_t s);
hipError_t launch_particles_advance(const ParticleArgs& a, hipStream_t s);
// Zero the drag-field footprint of the particles at their current positions.
hipError_t launch_particles_clear_footprint(const ParticleArgs& a, hipStream_t s);
// drag = alpha drag_new + (1 - alpha) drag_old; drag_old = drag (active particles).
hipError_t launch_particles_relax(const ParticleArgs& a, double alpha, hipStream_t s);
// p[0 .. n) = 0 in the real type (double when dbl).
hipError_t launch_zero_real(bool dbl, void* p, size_t n, hipStream_t s);
// force[d][c] += add[d][c] on cells whose kind is not solid (every cell
// without flags), d = 0..2, c < N.
hipError_t launch_add_field_fluid(bool dbl, void* force, const void* add, const uint32_t* flags, uint32_t N,
                                  hipStream_t s);
hipError_t launch_particle_index(const float* pos, int32_t n, int32_t nx, int32_t ny, int32_t nz, int32_t* idx,
                                 float* frac, hipStream_t s);

struct AdjointArgs {
  StepArgs fw;          // forward description; fw.fin = forward state f^n
  const void* lam_in;   // adjoint of f^{n+1} (post-collision), [19][P]
  void* lam_out;        // adjoint of f^n
  double* grad_perm;    // [N]
  double* partial;      // [nblk][4]: dJ/du_in x,y,z and J
  int objective;
};
hipError_t launch_adjoint(int storage, const AdjointArgs& a, const uint32_t* rflags, const Grid& g, hipStream_t s);
hipError_t launch_adjoint_final(const double* partial, uint32_t nblk, double* acc, hipStream_t s);

}  // namespace pol
