/*
 * pourover_hip.h -- C ABI of the MI355X (gfx950) D3Q19 lattice-Boltzmann engine.
 *
 * The library (libpourover_hip.so) owns no field memory: the caller allocates
 * every device buffer (PyTorch-ROCm tensors in the Python host layer) and
 * passes raw device pointers, which the library borrows and never frees.  A
 * context owns only its descriptor and a small scratch area (reduction
 * partials) allocated with hipMalloc in pol_create and released in
 * pol_destroy.
 *
 * Every entry point returns an int status (POL_OK == 0).  On failure the
 * context keeps a message readable with pol_last_error().  Work is enqueued
 * asynchronously on the hipStream_t passed as `stream` (NULL = legacy
 * default stream); only entry points that hand results back to host memory
 * (pol_reduce, pol_macroscopic / pol_step_ex / pol_step_from with a
 * reduction, pol_adjoint_collect) synchronise that stream.
 *
 * The reference (latteine1217/pour-over-coffee-lbm) is Python + Taichi and
 * has no C ABI; each entry point below names the reference interface whose
 * job it takes over.
 *
 * Population layout in HBM: f[q][z][y][x], q in the D3Q19 order of
 * config/core.py:36-38, x fastest, two ping-pong buffers.  Buffers hold
 * post-collision populations; one step gathers (pull streaming) from one
 * buffer and writes the other.  Every storage format (F32, F16, BF16, F64)
 * holds the scaled deviation D_q = 36 (f_q - w_q) from the lattice weight
 * (36 w_q = 12, 2, 1 are exact, so no rounded weight enters the arithmetic).
 */
#ifndef POUROVER_HIP_H
#define POUROVER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define POL_ABI_VERSION 5
#define POL_Q 19
#define POL_NRED 12 /* length of the reduction vector, see POL_RED_* */

/* status codes; mapped on the Python side onto the reference's BackendError
 * hierarchy (src/core/backends/compute_backends.py:35-65) */
enum {
  POL_OK = 0,
  POL_ERR_INVALID = 1,     /* bad argument / descriptor -> ComputeExecutionError */
  POL_ERR_HIP = 2,         /* HIP runtime error        -> ComputeExecutionError */
  POL_ERR_ALLOC = 3,       /* scratch allocation       -> MemoryAllocationError */
  POL_ERR_NO_DEVICE = 4,   /* no usable GPU            -> PLATFORM_UNAVAILABLE  */
  POL_ERR_STATE = 5,       /* call order (e.g. not bound) */
  POL_ERR_UNSUPPORTED = 6
};

/* population storage formats */
enum { POL_F32 = 0, POL_F16 = 1, POL_BF16 = 2, POL_F64 = 3 };

/* pol_desc.options */
#define POL_PERIODIC_X 0x1u
#define POL_PERIODIC_Y 0x2u
#define POL_PERIODIC_Z 0x4u
#define POL_USE_FLAGS 0x8u      /* cell_flags buffer drives walls / cell kinds */
#define POL_LES 0x10u           /* Smagorinsky eddy viscosity */
#define POL_FORCE_FIELD 0x20u   /* add per-cell force density buffer */
#define POL_PERM_FIELD 0x40u    /* per-cell permeability buffer (else darcy_K) */
#define POL_GRAVITY 0x80u       /* constant body-force density gravity[3] */
#define POL_PARTICLE_FIELD 0x100u /* implicit particle-drag field (pol_buffers.particle_drag) */
#define POL_LES_FD 0x200u       /* the eddy viscosity of POL_LES (while it is on) from the reference's
                                   finite differences (les_turbulence.py:318-380) instead of the
                                   local-stress form, read from pol_buffers.nu_sgs (pol_les_viscosity) */

/* cell flag word (uint32 per cell) */
#define POL_LINK_MASK 0x7FFFFu  /* bit q: source x - e_q of link q is a wall -> halfway bounce-back */
#define POL_KIND_SHIFT 24
#define POL_KIND_MASK 0x7u
#define POL_KIND_FLUID 0u
#define POL_KIND_SOLID 1u
#define POL_KIND_INLET 2u       /* equilibrium velocity inlet (inlet_rho, inlet_u) */
#define POL_KIND_OUTLET 3u      /* pressure outlet: f* = feq(outlet_rho, u) */
#define POL_FLAG_POROUS 0x08000000u /* Darcy-Forchheimer resistance, LES off */
#define POL_FLAG_NO_LES 0x10000000u

/* pol_step_ex / pol_macroscopic `what` bits */
#define POL_WANT_MACRO 0x1u   /* write rho/u buffers */
#define POL_WANT_REDUCE 0x2u  /* FP64 reductions into red_out[POL_NRED] */

/* reduction vector slots (FP64) */
#define POL_RED_MASS 0      /* sum rho over non-solid cells */
#define POL_RED_MOMX 1      /* sum rho*u */
#define POL_RED_MOMY 2
#define POL_RED_MOMZ 3
#define POL_RED_KE 4        /* sum rho*|u|^2/2 */
#define POL_RED_UMAX 5      /* max |u| */
#define POL_RED_NONFINITE 6 /* count of cells with non-finite rho or u */
#define POL_RED_OUTFLUX 7   /* sum over outlet cells of -rho*u_z */
#define POL_RED_CELLS 8     /* number of non-solid cells */
#define POL_RED_RHOMIN 9
#define POL_RED_RHOMAX 10
#define POL_RED_INFLUX 11   /* sum over inlet cells of -rho*u_z */

typedef struct pol_ctx pol_ctx;

typedef struct pol_desc {
  int32_t nx, ny, nz;
  int32_t storage;     /* POL_F32 / POL_F16 / POL_BF16 / POL_F64 */
  uint32_t options;    /* POL_PERIODIC_* | POL_USE_FLAGS | ... */
  int32_t device;      /* HIP device ordinal the buffers live on */
  double tau;          /* molecular relaxation time (nu = (tau - 1/2)/3) */
  double cs_smag;      /* Smagorinsky constant C_s (filter width 1 lattice unit) */
  double nu_sgs_max;   /* clamp on the eddy viscosity */
  double gravity[3];   /* constant body-force density (POL_GRAVITY) */
  double darcy_K;      /* uniform permeability in porous cells */
  double forch_CF;     /* Forchheimer coefficient C_F */
  double inlet_rho;
  double inlet_u[3];
  double outlet_rho;
} pol_desc;

/* Device buffers.  "real" = double for POL_F64 storage, float otherwise. */
typedef struct pol_buffers {
  void *f[2];           /* [19][nz][ny][nx] storage type, ping-pong */
  uint32_t *cell_flags; /* [nz][ny][nx] (POL_USE_FLAGS) */
  void *force;          /* real [3][cells] force density (POL_FORCE_FIELD) */
  void *perm;           /* real [cells] permeability (POL_PERM_FIELD) */
  void *rho;            /* real [cells] macroscopic output, may be NULL */
  void *u;              /* real [3][cells] macroscopic output, may be NULL */
  const uint32_t *active_chunks; /* dense layout: first-cell indices of the
                                    chunk_cells-cell row chunks (x0 %
                                    chunk_cells == 0) that hold at least one
                                    non-solid cell, or NULL to launch over
                                    every cell.  With a list the kernels touch
                                    only those chunks; their solid lanes store
                                    zero deviations (f = w), so written lines
                                    are whole.  Compact layout (required
                                    there): the first-cell indices of the
                                    launched row pieces (below) */
  uint32_t n_chunks;
  uint32_t chunk_cells; /* cells per chunk: 4, 8, 16, 32 or 64 (0 = POL_CHUNK) */
  void *particle_drag; /* real [4][cells] (POL_PARTICLE_FIELD): per cell
                          sum_p w beta_p v_p (x, y, z) and B = sum_p w beta_p
                          over the particles whose trilinear footprint covers
                          the cell (w = trilinear weight, beta = |drag|/|u_rel|).
                          The collision treats the drag B (V - u) implicitly in
                          u, like the Darcy term, so any B is stable. */
  /* Compact row-run population layout (NULL = dense [19][nz][ny][nx]).
   * Population q of cell (x, y, z) lives at f[q * stored_cells + slot],
   * slot = row_base(y, z) + x, for the cells of the row's stored run [x_begin,
   * x_end) (first to last non-solid cell of the row, possibly padded); rows
   * follow one another in (z, y) order.  Slots 0, 1, stored_cells - 2 and
   * stored_cells - 1 hold no cell (guards: the kernels clamp a row offset
   * into [1, stored_cells - 2] and reach x +- 1 by constant offsets).
   * Launches go over the chunk_cells-cell row pieces listed in
   * active_chunks (n_chunks piece words: the first-cell index, plus
   * POL_PIECE_PLAIN when every cell of the piece with x < nx has the flag
   * word 0, which the kernels then take without loading it; so
   * nx ny nz < 2^31) and read, for a piece's row
   * (z ny + y = first cell / nx), its record of POL_RECORD_WORDS uint32 in
   * chunk_records, which holds one record per dense row (nz * ny records,
   * 16-byte aligned; rows without a run are never read):
   *   [0] dense index of the row's first cell, [1 + a + 3 b] row_base of
   *   row (y + a - 1, z + b - 1) (a, b in 0..2; a row past a non-periodic
   *   face is the row itself, periodic faces wrap), [10] the row's run
   *   x_begin | x_end << 16, [11] the row's y | z << 16 (so nx, ny,
   *   nz < 65536).
   * Every stored slot is written each step (solid cells of a run with zero
   * deviations). */
  const uint32_t *chunk_records;
  uint32_t stored_cells; /* plane stride of the compact layout */
  uint32_t layout_flags; /* POL_LAYOUT_* */
  void *nu_sgs;          /* real [cells] eddy viscosity (POL_LES_FD): written by
                            pol_les_viscosity, read by the steps */
  const void *les_phase; /* real [cells] phase field of the |phi| < 0.9 interface
                            suppression (POL_LES_FD), or NULL (phi = 1) */
} pol_buffers;
/* No run of the compact layout holds a cell at x = 0 or x = nx - 1: the
 * kernels then leave out the x-face wrap / clamp of the gather.  Set it only
 * when true (otherwise x-face cells read wrong, in-bounds sources). */
#define POL_LAYOUT_NO_X_FACE 1u
#define POL_CHUNK 32 /* default chunk_cells */
#define POL_RECORD_WORDS 12 /* uint32 per compact-layout row record */
/* Compact-layout piece word bit: every cell of the piece inside its row has
 * cell_flags 0 (fluid, no bounce link, not porous, LES on).  Set it only
 * when true: the kernels use flag word 0 for those cells. */
#define POL_PIECE_PLAIN 0x80000000u

/* ---- library ------------------------------------------------------------ */
/* ABI version (POL_ABI_VERSION). */
int pol_version(void);
/* Static message for a status code. */
const char *pol_status_string(int status);
/* Number of HIP devices.  Replaces PlatformDetector platform probing
 * (src/core/backends/compute_backends.py:121-329). */
int pol_device_count(int *count);

/* ---- context lifecycle --------------------------------------------------- */
/* Create a solver context.  Replaces ComputeBackend.initialize_backend
 * (compute_backends.py:429-441) and the solver construction of
 * UnifiedLBMSolver (src/core/lbm_unified.py:44-66). */
int pol_create(pol_ctx **out, const pol_desc *desc);
/* Release the context and its scratch.  Replaces cleanup_backend
 * (compute_backends.py:443-452). */
int pol_destroy(pol_ctx *ctx);
/* Last error message of this context ("" if none); with ctx NULL, the message
 * of the last failed context-free call (pol_reduce_noalloc) on this thread. */
const char *pol_last_error(const pol_ctx *ctx);
/* Attach caller-owned device buffers.  Replaces the memory-adapter attribute
 * contract f/f_new/rho/u/solid (src/core/backends/cpu_backend.py:207-212). */
int pol_bind(pol_ctx *ctx, const pol_buffers *bufs);
/* Replace the physics scalars (tau, cs_smag, nu_sgs_max, gravity, darcy_K,
 * forch_CF, inlet/outlet values) of a context; nx/ny/nz, storage and device
 * must equal the context's, and so must the options except POL_LES, which
 * may be switched (LBMSolverProtocol.enable_les_turbulence,
 * src/core/lbm_protocol.py:168).  Takes effect for launches enqueued
 * afterwards.  Replaces setting config module constants between runs
 * (config/core.py:57-93, config/physics.py:255-273). */
int pol_set_params(pol_ctx *ctx, const pol_desc *desc);
/* Index (0/1) of the buffer holding the current state. */
int pol_current(const pol_ctx *ctx, int *index);
/* Select which buffer holds the current state (after a host-side restore). */
int pol_set_current(pol_ctx *ctx, int index);

/* ---- hot path ------------------------------------------------------------ */
/* f[current] = feq(rho, u) per cell (rho NULL -> 1, u NULL -> 0), computed in
 * the real type.  Replaces LBMSolver.init_fields
 * (src/core/legacy/lbm_solver.py:1067-1112). */
int pol_init_equilibrium(pol_ctx *ctx, const void *rho, const void *u, void *stream);
/* nsteps fused stream-collide passes.  Replaces
 * ComputeBackend.execute_collision_streaming (compute_backends.py:367-378,
 * cpu_backend.py:192-237) and LBMSolver._collision_streaming_step +
 * swap_fields (legacy/lbm_solver.py:460-487,630-654). */
int pol_step(pol_ctx *ctx, int nsteps, void *stream);
/* As pol_step; the last step also writes rho/u and/or fills red_out (FP64,
 * POL_NRED values, moments of the state that step collided). */
int pol_step_ex(pol_ctx *ctx, int nsteps, uint32_t what, double *red_out, void *stream);
/* One step from an explicit buffer `fin` into `fout` (same layout/storage as
 * the bound f); the context's current-buffer index is left unchanged.  Used
 * to record forward states for the adjoint's checkpoint recomputation. */
int pol_step_from(pol_ctx *ctx, const void *fin, void *fout, uint32_t what, double *red_out, void *stream);
/* nsteps collisionless streaming passes (f[next] = the populations a step
 * would gather -- bounce-back links, open faces, the layout's stores -- with
 * no collision and no moments), swapping the current buffer like pol_step.
 * Free streaming of the populations (omega = 0), and the step's memory access
 * pattern without its arithmetic: timing it against pol_step gives the
 * layout's bandwidth ceiling for the step.  No counterpart in the reference
 * (its streaming always follows a collision). */
int pol_stream_only(pol_ctx *ctx, int nsteps, void *stream);
/* Moments of the current state (the gather the next step would collide):
 * rho/u buffers and/or FP64 reductions.  Replaces
 * compute_macroscopic_quantities (compute_backends.py:393-404,
 * legacy/lbm_solver.py:488-535) and the statistics reductions of
 * visualizer.py:120-167 / numerical_stability.py:52-110. */
int pol_macroscopic(pol_ctx *ctx, uint32_t what, double *red_out, void *stream);
/* FP64 reductions of the current state; synchronous.  = pol_macroscopic
 * with POL_WANT_REDUCE. */
int pol_reduce(pol_ctx *ctx, double *red_out, void *stream);

/* ---- finite-difference eddy viscosity (src/physics/les_turbulence.py) ---- */
/* nu_sgs = (C_s * 1)^2 |S|, |S| = sqrt(2 S:S), S from central differences of
 * u_field (real [3][cells], NULL: the bound u buffer) over the interior
 * cells, in float32 in the reference's operation order
 * (_compute_sgs_from_vector, les_turbulence.py:318-380): zero where the cell
 * flags switch LES off (POL_FLAG_NO_LES, the filter paper's mask), where
 * |S| < 1e-3, where |phi| < 0.9 (les_phase), on the box faces, and clamped to
 * nu_sgs_max (the reference's 0.1).  C_s is the descriptor's cs_smag.
 * Needs POL_LES | POL_LES_FD.  With POL_LES_FD every pol_step_ex /
 * pol_coupled_step step first forms nu_sgs from the bound u and then writes
 * u (the velocity the legacy solver's next step reads,
 * legacy/lbm_solver.py:856-857), and the collision uses
 * tau = clamp(tau + 3 nu_sgs, 0.55, 1.90) (legacy/lbm_solver.py:565-572). */
int pol_les_viscosity(pol_ctx *ctx, const void *u_field, void *stream);

/* ---- allocation-free reduction (no counterpart in the reference) ---------- */
/* For callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc: the FP64
 * reductions of pol_reduce (POL_RED_* slots) of a state described without a
 * context.  desc and bufs as for pol_create / pol_bind; the state is
 * bufs->f[0] (f[1] unused); red_out_dev: POL_NRED doubles of device memory,
 * written asynchronously on `stream`.  Calls no hipMalloc / hipFree: the
 * per-block partials live in a static device array of
 * POL_NOALLOC_MAX_BLOCKS x POL_NRED doubles (6 MiB, which each GPU the
 * library runs on holds until the process ends), so a state of more than
 * POL_NOALLOC_MAX_BLOCKS x 256 work items (cells, or listed chunk lanes) is
 * rejected with POL_ERR_INVALID.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (concurrent streams or threads) race and
 * must be serialised by the caller.  On failure pol_last_error(NULL) gives
 * the message. */
#define POL_NOALLOC_MAX_BLOCKS 65536
int pol_reduce_noalloc(const pol_desc *desc, const pol_buffers *bufs, double *red_out_dev, void *stream);

/* ---- Lagrangian coffee particles (src/physics/coffee_particles.py) ------- */
/* drag magnitude laws (C_D from the Schiller-Naumann piecewise form of
 * coffee_particles.py:1088-1099 in both):
 *   TEXTBOOK : |F| = 1/2 rho C_D pi r^2 |u_rel|^2  (docs/mathematical/mathematical_models.md:250-275;
 *              Stokes 3 pi mu d |u_rel| at low Re)
 *   REFERENCE: |F| = 1/2 rho C_D pi r^2 |u_rel|    (coffee_particles.py:1137-1145 as written) */
#define POL_DRAG_TEXTBOOK 0
#define POL_DRAG_REFERENCE 1

typedef struct pol_particle_desc {
  int32_t count;         /* number of particle slots */
  int32_t drag_law;      /* POL_DRAG_TEXTBOOK or POL_DRAG_REFERENCE */
  double rho_fluid;      /* fluid density in drag law */
  double nu_fluid;       /* kinematic viscosity (mu = rho_fluid * nu_fluid) */
  double drag_max_per_mass; /* |drag| <= this * mass */
  double relax_alpha;    /* under-relaxation factor */
  double gravity[3];     /* acceleration on particles */
  double dt;             /* particle time step (lattice units) */
  double max_disp;       /* displacement clamp per step */
  double wall_damping;   /* velocity factor when a move is rejected */
  int32_t reaction;      /* POL_REACTION_IMPLICIT or POL_REACTION_EXPLICIT */
  int32_t reserved;      /* 0 */
} pol_particle_desc;

/* How the particles act back on the fluid (pol_particle_desc.reaction).
 *   IMPLICIT: the coupling spreads (w beta v_p, w beta) into the lattice's
 *             particle-drag field (pol_buffers.particle_drag, POL_PARTICLE_FIELD)
 *             and the collision applies sum w beta (v_p - u) with the cell
 *             velocity u implicit (stable for any drag coefficient).
 *   EXPLICIT: the reference's form: the coupling spreads -drag_new into a
 *             [3][cells] reaction field (pol_particle_buffers.reaction_field,
 *             coffee_particles.py:1146-1149,1049-1086), which
 *             pol_add_reaction_forces adds to the body force on the fluid
 *             cells (legacy/lbm_solver.py:1478-1483). */
#define POL_REACTION_IMPLICIT 0
#define POL_REACTION_EXPLICIT 1

/* All vectors SoA [3][count], in the real type of the context (double for
 * POL_F64 storage, float otherwise); active is int32. */
typedef struct pol_particle_buffers {
  void *pos, *vel, *radius, *mass;
  void *active;
  void *drag_new, *drag_old, *drag; /* [3][count] */
  void *u_fluid;  /* [3][count] interpolated fluid velocity (diagnostic) */
  void *re_p, *cd; /* [count] diagnostics */
  void *drag_field; /* [4][cells] particle-drag field (= pol_buffers.particle_drag); IMPLICIT */
  const void *u_field; /* [3][cells] fluid velocity field read by the gather */
  void *reaction_field; /* [3][cells] reaction field (real type); EXPLICIT, else NULL */
} pol_particle_buffers;

/* Attach a particle system.  Replaces CoffeeParticleSystem field set-up
 * (coffee_particles.py:30-64). */
int pol_particles_bind(pol_ctx *ctx, const pol_particle_desc *d, const pol_particle_buffers *b);
/* Interpolate, drag, under-relax and spread the reaction: IMPLICIT adds
 * (w beta v_p, w beta) to the particle-drag field, EXPLICIT clears the
 * reaction field and adds w (-drag_new) to it.  A second call before
 * pol_particles_advance recomputes (IMPLICIT clears its own earlier
 * footprint first), so the field never holds two couplings.  Replaces
 * compute_two_way_coupling_forces + apply_under_relaxation
 * (coffee_particles.py:1107-1154,1200-1212). */
int pol_particles_couple(pol_ctx *ctx, void *stream);
/* compute_two_way_coupling_forces(fluid_u) (coffee_particles.py:1107-1154)
 * alone: interpolate u_field (NULL: the bound u_field), drag_new, Reynolds
 * number and C_D per particle, and the reaction as in pol_particles_couple,
 * without the under-relaxation. */
int pol_particles_compute_coupling(pol_ctx *ctx, const void *u_field, void *stream);
/* apply_under_relaxation(alpha) (coffee_particles.py:1200-1212): on active
 * particles drag = alpha drag_new + (1 - alpha) drag_old; drag_old = drag.
 * alpha in [0, 1]. */
int pol_particles_relax(pol_ctx *ctx, double alpha, void *stream);
/* clear_reaction_forces (coffee_particles.py:1101-1105): EXPLICIT zeroes the
 * reaction field; IMPLICIT zeroes the drag-field footprint of the last
 * coupling if it is still there. */
int pol_particles_clear_reaction(pol_ctx *ctx, void *stream);
/* add_particle_reaction_forces (legacy/lbm_solver.py:1478-1483): the body
 * force (pol_buffers.force, POL_FORCE_FIELD) takes reaction_field ([3][cells],
 * real type; NULL: the bound particle system's reaction field) on every
 * cell whose kind is not solid. */
int pol_add_reaction_forces(pol_ctx *ctx, const void *reaction_field, void *stream);
/* Clear this step's particle-drag footprint and advance the particles.
 * Replaces clear_reaction_forces + update_particle_physics
 * (coffee_particles.py:1101-1105,641-720). */
int pol_particles_advance(pol_ctx *ctx, void *stream);
/* nsteps two-way coupled steps, each: pol_particles_couple, (EXPLICIT: the
 * body force is cleared and takes the reaction field, as
 * step_with_two_way_coupling clears and refills it), one fused
 * stream-collide step writing rho/u (the velocity the next coupling
 * interpolates), pol_particles_advance.  The last step also fills red_out
 * with `what & POL_WANT_REDUCE` (synchronous then).  Replaces the per-step
 * loop of LBMSolver.step_with_two_way_coupling (legacy/lbm_solver.py:1485-1509)
 * and main.py's coupling order (main.py:795-839). */
int pol_coupled_step(pol_ctx *ctx, int nsteps, uint32_t what, double *red_out, void *stream);
/* Particle -> cell index and clamped fractions, the formula of
 * coffee_particles.py:1159-1174: i = (int)max(0, min(n-2, p)),
 * frac = clamp(p - i, 0, 1).  pos SoA [3][n]; idx [3][n]; frac [3][n]. */
int pol_particle_index(const float *pos, int32_t n, int32_t nx, int32_t ny, int32_t nz,
                       int32_t *idx, float *frac, void *stream);

/* ---- adjoint (no counterpart in the reference) ---------------------------- */
#define POL_RFLAG_CLAMP 0x80000000u /* rev_flags bit 31: cell lies on a non-periodic face
                                       (clamped readers of its populations possible) */

typedef struct pol_adjoint_buffers {
  void *lam[2];        /* adjoint populations [19][cells], real type */
  double *grad_perm;   /* [cells] accumulated dJ/dK (FP64), may be NULL */
  uint32_t *rev_flags; /* [cells] reverse-gather word: bit p = the cell
                          x + e_p gathers population p of x normally;
                          bit 31 = clamped readers possible (open face) */
} pol_adjoint_buffers;
/* Attach adjoint buffers (lam[0] is the current adjoint state).  F32 and F64
 * storage only. */
int pol_adjoint_bind(pol_ctx *ctx, const pol_adjoint_buffers *b);
/* One reverse step: given the forward state `f_state` (the buffer the forward
 * step gathered from, same layout/storage as f) and the current adjoint, write
 * the adjoint of f_state.  With objective != 0 the step's outlet mass flux
 * (sum over outlet cells of -rho*u_z, POL_RED_OUTFLUX) is added to J.
 * dJ/dK accumulates in grad_perm, dJ/d(inlet_u) and J in the context (FP64).
 * Steps must be issued in reverse time order. */
int pol_adjoint_step(pol_ctx *ctx, const void *f_state, int objective, void *stream);
/* count reverse steps in one call: pol_adjoint_step with the forward states
 * states + i * stride_bytes for i = count - 1 down to 0, the objective on for
 * i >= objective_first. */
int pol_adjoint_sweep(pol_ctx *ctx, const void *states, uint64_t stride_bytes, int count, int objective_first,
                      void *stream);
/* count - 1 fused steps in one call, each from states + (i - 1) * stride_bytes
 * into states + i * stride_bytes (i = 1 .. count - 1): rebuilds a run of
 * forward states from the first (the checkpointed reverse sweep; no
 * counterpart in the reference). */
int pol_step_chain(pol_ctx *ctx, void *states, uint64_t stride_bytes, int count, void *stream);
/* Zero the adjoint state, grad_perm and the inlet-velocity gradient. */
int pol_adjoint_reset(pol_ctx *ctx, void *stream);
/* Fold and copy the FP64 accumulators {dJ/du_in x, y, z, J} (4 doubles) to
 * host; synchronous. */
int pol_adjoint_collect(pol_ctx *ctx, double *out4, void *stream);

/* ---- field diagnostics (src/core/numerical_stability.py,
 *      src/visualization/lbm_diagnostics.py, src/visualization/visualizer.py) --
 * The reference reduces host copies of rho, u and f in numpy / Taichi loops
 * (numerical_stability.py:53-110, lbm_diagnostics.py:188-460,
 * visualizer.py:120-167).  Here one device pass over the bound rho / u
 * buffers (and optionally the populations) produces every sum, count and
 * extreme those consumers read, in FP64 and in a fixed block order
 * (deterministic).  Magnitudes |u| are formed in FP32 as
 * sqrt((ux*ux + uy*uy) + uz*uz) without contraction, the per-element value
 * the reference's float32 numpy expressions give. */
#define POL_STAT_CELLS 0          /* all cells */
#define POL_STAT_SUM_DRHO 1       /* sum (rho - 1) */
#define POL_STAT_SUM_DRHO2 2      /* sum (rho - 1)^2 */
#define POL_STAT_MAX_ABS_DRHO 3   /* max |rho - 1| (NaN if any rho is NaN) */
#define POL_STAT_MOM_X 4          /* sum rho u */
#define POL_STAT_MOM_Y 5
#define POL_STAT_MOM_Z 6
#define POL_STAT_SUM_UMAG 7       /* sum |u| */
#define POL_STAT_SUM_UMAG2 8      /* sum |u|^2 */
#define POL_STAT_MAX_UMAG 9       /* max |u| (NaN if any |u| is NaN) */
#define POL_STAT_SUM_MACH_ACTIVE 10 /* sum of Mach = |u| / c_s (FP32) over Mach > 1e-10 */
#define POL_STAT_ACTIVE 11        /* cells with Mach > 1e-10 */
#define POL_STAT_MACH_GT 12       /* cells with Mach > 0.1 */
#define POL_STAT_SUM_UX 13
#define POL_STAT_SUM_UY 14
#define POL_STAT_SUM_UZ 15
#define POL_STAT_SUM_UX2 16
#define POL_STAT_SUM_UY2 17
#define POL_STAT_SUM_UZ2 18
#define POL_STAT_SUM_UMAG_TOP 19  /* sum |u| over z >= z_top */
#define POL_STAT_SUM_UMAG_MID 20  /* z_mid <= z < z_top */
#define POL_STAT_SUM_UMAG_BOT 21  /* z < z_mid */
#define POL_STAT_OUTLET_DOWN 22   /* sum over z < z_outlet of max(-u_z, 0) */
#define POL_STAT_DRHO_PREV2 23    /* sum (rho - prev_rho)^2 (prev_rho given) */
#define POL_STAT_DU_PREV2 24      /* sum |u - prev_u|^2 (prev_u given) */
#define POL_STAT_FLUID 25         /* non-solid cells */
#define POL_STAT_FLUID_RHO_NAN 26
#define POL_STAT_FLUID_RHO_INF 27
#define POL_STAT_FLUID_RHO_MIN 28 /* over finite rho; 1e300 when none */
#define POL_STAT_FLUID_RHO_MAX 29 /* over finite rho; -1e300 when none */
#define POL_STAT_FLUID_U_NAN 30   /* |u| NaN */
#define POL_STAT_FLUID_U_INF 31   /* |u| infinite */
#define POL_STAT_FLUID_UMAX 32    /* over finite |u|; 0 when none */
#define POL_STAT_FLUID_POP_NAN 33 /* NaN populations (with_populations) */
#define POL_STAT_FLUID_POP_INF 34 /* infinite populations (with_populations) */
#define POL_STAT_INT_CELLS 35     /* non-solid cells with 1 <= x, y, z <= n - 2 */
#define POL_STAT_INT_WATER 36     /* sum rho over those with phase > 0.1 */
#define POL_STAT_INT_AIR 37       /* sum rho over the others */
#define POL_STAT_INT_UMAX 38      /* max |u| over interior non-solid cells (NaN |u| skipped); -1e300 when none */
#define POL_STAT_INT_UMIN 39      /* min |u| (NaN |u| skipped); 1e300 when none */
#define POL_STAT_INT_USUM 40      /* sum |u| */
#define POL_STAT_SUM_UMAG_ACTIVE 41 /* sum |u| over |u| > 1e-10 */
#define POL_STAT_UMAG_ACTIVE 42   /* cells with |u| > 1e-10 */
#define POL_NSTAT 43

typedef struct pol_stats_opts {
  const void *prev_rho;    /* real [cells] previous density, or NULL */
  const void *prev_u;      /* real [3][cells] previous velocity, or NULL */
  const void *phase;       /* real [cells] phase field, or NULL (every cell water) */
  int32_t z_top, z_mid;    /* band edges of the regional |u| sums */
  int32_t z_outlet;        /* outlet slab: z < z_outlet */
  int32_t with_populations; /* also count NaN / Inf populations of the current state */
} pol_stats_opts;
/* Statistics of the bound rho / u buffers (compute them first, e.g. with
 * pol_macroscopic) into out[POL_NSTAT] (host) and, when z_profile is not
 * NULL, sum_{x,y} rho per z plane into z_profile[nz] (host).  Synchronous. */
int pol_field_stats(pol_ctx *ctx, const pol_stats_opts *opts, double *out, double *z_profile, void *stream);
/* magnitude[cell] = |u| (FP32 form above), pressure[cell] = rho * (float)(1/3)
 * (the export fields of enhanced_visualizer.py:1487-1505); real [cells]
 * device buffers. */
int pol_export_fields(pol_ctx *ctx, void *magnitude, void *pressure, void *stream);
/* Emergency stabilisation of the non-solid cells
 * (numerical_stability.py:165-197): rho NaN / Inf / <= 0 -> 1, else clamped
 * into [min_density, max_density]; |u| NaN / Inf -> u = 0, |u| > max_velocity
 * -> scaled to max_velocity; every NaN / Inf population of the current state
 * -> the equilibrium of the repaired (rho, u). */
int pol_stabilize(pol_ctx *ctx, double max_velocity, double min_density, double max_density, void *stream);

/* ---- thermal D3Q7 (src/physics/thermal_lbm.py) -----------------------------
 * Temperature populations g_q (q = rest, +x, -x, +y, -y, +z, -z; weights 1/4,
 * 1/8 x 6) over every cell of the box, stored after streaming like the
 * reference's `g`, equilibrium w_q T.  One step (thermal_lbm.py:396-427):
 *   T = sum_q g_q; S = heat_source (+ -u.grad T by central differences on
 *   interior cells, added into heat_source, with POL_THERMAL_CONVECTION);
 *   g*_q = (g_q - omega (g_q - w_q T)) + w_q S; streamed to x + e_q, a
 *   population whose source x - e_q lies outside the box keeping the cell's
 *   own g*_q (thermal_lbm.py:231-253); T' = sum_q g'_q.
 * FP32 arithmetic in the reference's operation order without contraction
 * (FP64 for POL_F64 contexts). */
#define POL_THERMAL_CONVECTION 0x1u
/* Every step starts from heat_source_base: S = base (+ convection), written
 * to heat_source -- the coupled loop's reset_heat_source_to_base before each
 * thermal step (thermal_fluid_coupled.py:193-194, one sub-cycle) fused into
 * the step. */
#define POL_THERMAL_RESET_SOURCE 0x2u

typedef struct pol_thermal_desc {
  double omega;        /* 1 / tau_thermal */
  uint32_t options;    /* POL_THERMAL_CONVECTION | POL_THERMAL_RESET_SOURCE */
  int32_t reserved;
} pol_thermal_desc;

typedef struct pol_thermal_buffers {
  void *g[2];            /* real [7][cells] streamed populations, ping-pong */
  void *temperature[2];  /* real [cells] T after the step / before it, ping-pong */
  void *heat_source;     /* real [cells] S (convection adds -u.grad T into it) */
  const void *velocity;  /* real [3][cells] advecting velocity, or NULL */
  const void *conductivity; /* real [cells] k of the heat flux, or NULL (0.68) */
  const void *heat_source_base; /* real [cells] base S for POL_THERMAL_RESET_SOURCE, or NULL */
} pol_thermal_buffers;
/* Attach the thermal buffers (g[0] / temperature[0] current). */
int pol_thermal_bind(pol_ctx *ctx, const pol_thermal_desc *d, const pol_thermal_buffers *b);
/* Change omega / options of a bound thermal state. */
int pol_thermal_set(pol_ctx *ctx, const pol_thermal_desc *d);
/* g_q = w_q T from the current temperature (thermal_lbm.py:132-170). */
int pol_thermal_init(pol_ctx *ctx, void *stream);
/* nsteps thermal steps (without convection the temperature of the
 * intermediate states is formed inside the next step's collide pass, so the
 * temperature pass runs once per call).  With unstable_out != NULL the last
 * step's stability test (thermal_lbm.py:348-376: T outside [-50, 150] or
 * NaN, a population NaN or |g| > 1000) is copied to *unstable_out (count of
 * failing cells); synchronous then. */
int pol_thermal_step(pol_ctx *ctx, int nsteps, int32_t *unstable_out, void *stream);
/* Heat flux q = -k grad T (central differences, interior cells; 0 on the
 * faces) of the current temperature into flux (real [3][cells];
 * thermal_lbm.py:268-286). */
int pol_thermal_heat_flux(pol_ctx *ctx, void *flux, void *stream);
/* {min, max, sum} of the current temperature (FP64, host); synchronous. */
int pol_thermal_stats(pol_ctx *ctx, double *out3, void *stream);
/* Boundary forms of thermal_lbm.py:288-346 on the cells where mask (uint8
 * [cells]) is non-zero: kind 1 Dirichlet (T = a, g = w a), kind 2 Neumann
 * (S = a), kind 3 convective (S = -a (T - b)). */
int pol_thermal_boundary(pol_ctx *ctx, int kind, const uint8_t *mask, double a, double b, void *stream);

/* ---- multiphase phase field (src/core/multiphase_3d.py) ---------------------
 * Cahn-Hilliard phase field phi (-1 air, +1 water) with CSF surface tension,
 * the producer of the rho / phase / body_force the stream-collide step reads
 * (main.py:795-800, 837-839).  Fields are float32 [cells] like the
 * reference's; rho, u and the body force are the bound context's buffers
 * (u as the last pol_macroscopic left it), phase is real [cells].  The
 * reference's step (multiphase_3d.py:389-414) never writes face cells:
 * their normal and surface force stay 0 and phi takes their phi_new values.
 * float32 arithmetic in the reference's operation order, no contraction. */
#define POL_MP_APPLY 0x1u /* pol_multiphase_interface: body_force += F / rho on fluid cells with rho > 1e-10 */

typedef struct pol_multiphase_desc {
  double sigma;           /* SURFACE_TENSION_COEFF (config.SURFACE_TENSION_LU) */
  double mobility;        /* MOBILITY M of M lap(mu) */
  double interface_width; /* INTERFACE_WIDTH W: kappa = 3 sigma W / 8 in the chemical potential */
  double rho_water;       /* config.RHO_WATER */
  double rho_air;         /* config.RHO_AIR */
  double dt;              /* config.DT */
} pol_multiphase_desc;

typedef struct pol_multiphase_buffers {
  float *phi[2];          /* [cells] phase field, ping-pong (phi[0] current after bind) */
  const float *phi_new;   /* [cells] the reference's phi_new: the step copies its face cells into phi */
  float *mu;              /* [cells] chemical potential */
  void *phase;            /* real [cells] lbm.phase, written by the step */
} pol_multiphase_buffers;

/* Optional outputs of pol_multiphase_interface (each may be NULL): the
 * reference's grad_phi, grad_mu, normal ([3][cells]), curvature ([cells])
 * and surface_force ([3][cells]) on interior cells. */
typedef struct pol_interface_fields {
  float *grad_phi;
  float *grad_mu;
  float *normal;
  float *curvature;
  float *surface_force;
} pol_interface_fields;

/* Attach the phase-field buffers to a bound context. */
int pol_multiphase_bind(pol_ctx *ctx, const pol_multiphase_desc *d, const pol_multiphase_buffers *b);
/* compute_chemical_potential (multiphase_3d.py:78-108): mu (and
 * laplacian_phi when not NULL) on interior cells of the current phi. */
int pol_multiphase_chemical_potential(pol_ctx *ctx, float *laplacian_phi, void *stream);
/* compute_gradients + compute_curvature + compute_surface_tension_force
 * (multiphase_3d.py:110-149, 365-381) of the current phi; with POL_MP_APPLY
 * also apply_surface_tension (:403-411) into the context's force buffer --
 * accumulate_surface_tension_pre_collision (:416-421) in one pass. */
int pol_multiphase_interface(pol_ctx *ctx, uint32_t flags, const pol_interface_fields *out, void *stream);
/* The rest of step (multiphase_3d.py:389-414): upwind advection with the
 * context's u, M lap(mu), clamp, phase separation, copy to phi (the
 * ping-pong flips) and update_density_from_phase into rho and phase. */
int pol_multiphase_step(pol_ctx *ctx, void *stream);
/* update_density_from_phase alone (multiphase_3d.py:413-428). */
int pol_multiphase_density(pol_ctx *ctx, void *stream);
/* Index (0 / 1) of the current phi buffer. */
int pol_multiphase_current(pol_ctx *ctx, int *index);

/* ---- body-force producers of the main loop ------------------------------------
 * Pouring (src/physics/precise_pouring.py), pressure drive
 * (src/physics/pressure_gradient_drive.py) and filter-paper dynamics
 * (src/physics/filter_paper.py) on the bound context's rho / body force /
 * cell flags (solid cells are skipped, as the reference skips solid == 1).
 * float32 arithmetic in the reference's operation order, no contraction. */
typedef struct pol_pour_desc {
  double center_x, center_y; /* pour position this step (lattice units) */
  double radius;             /* POUR_DIAMETER_GRID / 2 */
  double stream_height;      /* pour_stream_height (4.0) */
  double velocity;           /* POUR_VELOCITY (lu / ts) */
  double flow_rate;          /* pour_flow_rate */
  double dt;                 /* time step of this call */
  int32_t height;            /* POUR_HEIGHT: z of the nozzle */
  int32_t reserved;
} pol_pour_desc;
/* apply_pouring_force (precise_pouring.py:132-162): body_force.z -=
 * min(velocity I flow_rate / dt, 10) on fluid cells of the pour region,
 * I = exp(-0.5 (d / radius)^2) exp(-(height - z) / 2).  Only the region's
 * bounding box is launched. */
int pol_pour_force(pol_ctx *ctx, const pol_pour_desc *d, void *stream);
/* apply_gradual_phase_change (precise_pouring.py:164-193) on phi (float32
 * [cells]). */
int pol_pour_phase(pol_ctx *ctx, const pol_pour_desc *d, float *phi, void *stream);
/* apply_density_drive (pressure_gradient_drive.py:96-122; half != 0: the
 * mixed form :206-227): rho += limited(rate (target - rho)) clamped to
 * [lo, hi] on fluid cells; target is float32 [cells]. */
int pol_density_drive(pol_ctx *ctx, const float *target, double rate, int half, double max_adjust, double lo,
                      double hi, void *stream);
/* compute_pressure_gradient (pressure_gradient_drive.py:124-173) into
 * pressure_force (float32 [3][cells], fluid cells) and, mode 1, body_force +=
 * pressure_force (:186-190) or, mode 2, body_force += 0.5 pressure_force
 * (:276-279); mode 0 stores only.  pressure_force must hold zero at solid
 * cells, as the reference's field does (it starts at zero and only this pass
 * writes it); the pass may store that zero again.
 * mode | POL_PRESSURE_SURFACE_TENSION (modes 1 and 2, a bound phase field):
 * the pre-collision surface tension of pol_multiphase_interface(POL_MP_APPLY)
 * (multiphase_3d.py:403-421) follows in the same pass on the sum just formed
 * -- the result of the two calls in a row, bit for bit, with one read and one
 * write of the body force instead of two. */
#define POL_PRESSURE_SURFACE_TENSION 0x4
int pol_pressure_force(pol_ctx *ctx, float *pressure_force, int mode, double max_force, double cs2, void *stream);
/* update_dynamic_resistance (filter_paper.py:699-746) on the paper cells
 * listed in cells (ncells cell indices, the cells with filter_zone == 1);
 * accumulated / blockage float32 [cells of the grid]. */
int pol_filter_blockage(pol_ctx *ctx, const uint32_t *cells, int32_t ncells, float *accumulated, float *blockage,
                        void *stream);
/* block_particles_at_filter (filter_paper.py:616-697): float32 SoA
 * positions / velocities [3][count], int32 active; grid cell = int(pos /
 * scale).  The horizontal kick draws from a counter-based generator keyed
 * by seed and particle index (the reference uses ti.random). */
int pol_filter_block_particles(pol_ctx *ctx, const uint8_t *zone, float *accumulated, const float *pos, float *vel,
                               const int32_t *active, int32_t count, double scale, uint32_t seed, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* POUROVER_HIP_H */
