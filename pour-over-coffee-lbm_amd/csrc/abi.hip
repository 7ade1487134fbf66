// C ABI of libpourover_hip.so (declared in include/pourover_hip.h).
// Host-side only: validation, context state, launch dispatch.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "engine.h"

using namespace pol;

struct pol_ctx {
  pol_desc d;
  pol_buffers b;
  bool bound = false;
  int cur = 0;
  uint32_t N = 0;
  Grid grid{};   // dense grid over all cells (sizes the partial buffers)
  Grid wgrid{};  // work grid: over the active-cell list when bound, else = grid
  uint32_t cshift = 5;    // log2 of the cells per chunk of the bound list
  uint32_t P = 0;         // population plane stride (N, or the compact layout's stored cells)
  uint32_t part_cap = 0;  // blocks the partial buffers hold (any work grid fits)
  double* partial = nullptr;  // [nblk][POL_NRED]
  double* red_dev = nullptr;  // [POL_NRED]
  // particles
  bool particles = false;
  bool pcoupled = false;  // the implicit drag field holds the footprint of a coupling
  ParticleArgs pa{};
  // adjoint
  bool adjoint = false;
  pol_adjoint_buffers ab{};
  int acur = 0;
  double* adj_partial = nullptr;  // [nblk][4]
  double* adj_acc = nullptr;      // [4]: dJ/du_in (3), J
  // field statistics
  uint32_t stat_tiles = 0;          // blocks per z plane
  double* stat_partial = nullptr;   // [nz * stat_tiles][POL_NSTAT]
  double* stat_zpart = nullptr;     // [nz * stat_tiles]
  double* stat_dev = nullptr;       // [POL_NSTAT + nz]
  double* stat_host = nullptr;      // pinned copy of stat_dev (one transfer per call)
  // multiphase phase field
  bool mp = false;
  pol_multiphase_desc md{};
  pol_multiphase_buffers mb{};
  int mcur = 0;
  // thermal D3Q7
  bool thermal = false;
  pol_thermal_desc td{};
  pol_thermal_buffers tb{};
  int gcur = 0, tcur = 0;
  std::string err;
};

static thread_local std::string g_free_err;  // message of the last failed context-free call

static int fail(pol_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}
static int hip_fail(pol_ctx* c, hipError_t e, const char* where) {
  return fail(c, POL_ERR_HIP, std::string(where) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")");
}

static size_t storage_bytes(int s) {
  switch (s) {
    case POL_F32: return 4;
    case POL_F16: return 2;
    case POL_BF16: return 2;
    case POL_F64: return 8;
  }
  return 0;
}

extern "C" {

int pol_version(void) { return POL_ABI_VERSION; }

const char* pol_status_string(int status) {
  switch (status) {
    case POL_OK: return "ok";
    case POL_ERR_INVALID: return "invalid argument";
    case POL_ERR_HIP: return "HIP runtime error";
    case POL_ERR_ALLOC: return "device allocation failed";
    case POL_ERR_NO_DEVICE: return "no HIP device";
    case POL_ERR_STATE: return "invalid call order";
    case POL_ERR_UNSUPPORTED: return "unsupported configuration";
  }
  return "unknown status";
}

int pol_device_count(int* count) {
  if (!count) return POL_ERR_INVALID;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    *count = 0;
    return e == hipErrorNoDevice ? POL_ERR_NO_DEVICE : POL_ERR_HIP;
  }
  *count = n;
  return n > 0 ? POL_OK : POL_ERR_NO_DEVICE;
}

// Grid, storage and physics checks of a descriptor (no HIP call).
static bool desc_ok(const pol_desc& d) {
  if (d.nx < 2 || d.ny < 2 || d.nz < 2) return false;
  const uint64_t N = (uint64_t)d.nx * d.ny * d.nz;
  if (N >= (1ull << 31) || (uint64_t)d.nx * d.ny >= (1ull << 31)) return false;
  if (storage_bytes(d.storage) == 0) return false;
  if (!(d.tau > 0.5) || !std::isfinite(d.tau)) return false;
  if ((d.options & POL_LES) && !(d.cs_smag >= 0.0)) return false;
  if ((d.options & POL_LES_FD) && !(d.nu_sgs_max >= 0.0)) return false;
  return true;
}

int pol_create(pol_ctx** out, const pol_desc* desc) {
  if (!out || !desc) return POL_ERR_INVALID;
  *out = nullptr;
  const pol_desc& d = *desc;
  if (!desc_ok(d)) return POL_ERR_INVALID;
  const uint64_t N = (uint64_t)d.nx * d.ny * d.nz;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return POL_ERR_NO_DEVICE;
  if (d.device < 0 || d.device >= ndev) return POL_ERR_INVALID;
  pol_ctx* c = new (std::nothrow) pol_ctx();
  if (!c) return POL_ERR_ALLOC;
  c->d = d;
  c->N = (uint32_t)N;
  c->grid = make_grid(N);
  c->wgrid = c->grid;
  // the largest work grid: every row chunk listed, rows padded to 64 cells
  c->part_cap = make_grid((uint64_t)d.ny * d.nz * ((uint64_t)d.nx + 63)).nblk;
  hipError_t e = hipSetDevice(d.device);
  if (e == hipSuccess) e = hipMalloc((void**)&c->partial, (size_t)c->part_cap * POL_NRED * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&c->red_dev, POL_NRED * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&c->adj_partial, (size_t)c->part_cap * 4 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&c->adj_acc, 4 * sizeof(double));
  // statistics: up to 12 blocks per plane (2688 blocks at 224^3, each lane
  // ~8 groups of four cells: the per-block reduction of the 43 statistics is
  // amortised; 224^3 V60 kernel 0.062 ms against 0.069 / 0.067 ms with 8 / 24)
  const uint64_t nxy = (uint64_t)d.nx * d.ny;
  // 12 strips per plane; more for planes so large that a lane would take
  // 65536 cells (the statistics kernel counts in 16-bit halves)
  c->stat_tiles = (uint32_t)std::max<uint64_t>(std::min<uint64_t>((nxy + BLOCK - 1) / BLOCK, 12),
                                               (nxy + 128ull * 65000 - 1) / (128ull * 65000));
  const size_t sblk = (size_t)c->stat_tiles * d.nz;
  if (e == hipSuccess) e = hipMalloc((void**)&c->stat_partial, sblk * POL_NSTAT * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&c->stat_zpart, sblk * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&c->stat_dev, ((size_t)POL_NSTAT + d.nz) * sizeof(double));
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->stat_host, ((size_t)POL_NSTAT + d.nz) * sizeof(double));
  if (e != hipSuccess) {
    (void)hipHostFree(c->stat_host);
    (void)hipFree(c->partial);
    (void)hipFree(c->red_dev);
    (void)hipFree(c->adj_partial);
    (void)hipFree(c->adj_acc);
    (void)hipFree(c->stat_partial);
    (void)hipFree(c->stat_zpart);
    (void)hipFree(c->stat_dev);
    delete c;
    return POL_ERR_ALLOC;
  }
  *out = c;
  return POL_OK;
}

int pol_destroy(pol_ctx* c) {
  if (!c) return POL_OK;
  (void)hipSetDevice(c->d.device);
  (void)hipFree(c->partial);
  (void)hipFree(c->red_dev);
  (void)hipFree(c->adj_partial);
  (void)hipFree(c->adj_acc);
  (void)hipFree(c->stat_partial);
  (void)hipFree(c->stat_zpart);
  (void)hipFree(c->stat_dev);
  (void)hipHostFree(c->stat_host);
  delete c;
  return POL_OK;
}

const char* pol_last_error(const pol_ctx* c) {
  if (c) return c->err.c_str();
  return g_free_err.empty() ? "null context" : g_free_err.c_str();
}

// Checks the buffers against the context's descriptor and records them
// (no allocation, no HIP call).  pair: both population buffers required.
static int bind_buffers(pol_ctx* c, const pol_buffers* b, bool pair) {
  if (pair && (!b->f[0] || !b->f[1] || b->f[0] == b->f[1]))
    return fail(c, POL_ERR_INVALID, "pol_bind: two distinct population buffers required");
  if (!b->f[0]) return fail(c, POL_ERR_INVALID, "pol_bind: population buffer f[0] required");
  const uint32_t o = c->d.options;
  if ((o & POL_USE_FLAGS) && !b->cell_flags) return fail(c, POL_ERR_INVALID, "pol_bind: POL_USE_FLAGS without cell_flags");
  if ((o & POL_FORCE_FIELD) && !b->force) return fail(c, POL_ERR_INVALID, "pol_bind: POL_FORCE_FIELD without force buffer");
  if ((o & POL_PERM_FIELD) && !b->perm) return fail(c, POL_ERR_INVALID, "pol_bind: POL_PERM_FIELD without perm buffer");
  if ((o & POL_PARTICLE_FIELD) && !b->particle_drag)
    return fail(c, POL_ERR_INVALID, "pol_bind: POL_PARTICLE_FIELD without particle_drag buffer");
  if ((o & POL_LES_FD) && !b->nu_sgs) return fail(c, POL_ERR_INVALID, "pol_bind: POL_LES_FD without nu_sgs buffer");
  if ((o & POL_PERM_FIELD) && !(o & POL_USE_FLAGS)) return fail(c, POL_ERR_INVALID, "pol_bind: permeability needs cell flags");
  const uint32_t cc = b->chunk_cells ? b->chunk_cells : POL_CHUNK;
  if (cc != 4 && cc != 8 && cc != 16 && cc != 32 && cc != 64)
    return fail(c, POL_ERR_INVALID, "pol_bind: chunk_cells must be 4, 8, 16, 32 or 64");
  const uint32_t cshift = cc == 4 ? 2 : (cc == 8 ? 3 : (cc == 16 ? 4 : (cc == 32 ? 5 : 6)));
  const uint64_t max_chunks = (uint64_t)c->d.ny * c->d.nz * (((uint64_t)c->d.nx + cc - 1) / cc);
  // the dense layout may list chunks; the compact layout lists its pieces
  const bool listed = b->active_chunks != nullptr;
  if (b->chunk_records && !b->active_chunks)
    return fail(c, POL_ERR_INVALID, "pol_bind: the compact layout needs its piece list in active_chunks");
  if (listed && (b->n_chunks == 0 || b->n_chunks > max_chunks || ((uint64_t)b->n_chunks << cshift) >= (1ull << 32)))
    return fail(c, POL_ERR_INVALID, "pol_bind: n_chunks out of range");
  if (listed && !(o & POL_USE_FLAGS))
    return fail(c, POL_ERR_INVALID, "pol_bind: a chunk list needs cell flags (solid lanes are masked by them)");
  if (b->chunk_records && (uint64_t)c->d.nx * c->d.ny * c->d.nz >= (1ull << 31))
    return fail(c, POL_ERR_INVALID, "pol_bind: the compact layout needs nx ny nz < 2^31 (piece word bit 31 is POL_PIECE_PLAIN)");
  if (b->chunk_records && (b->stored_cells < 4 || b->stored_cells >= (1u << 31)))
    return fail(c, POL_ERR_INVALID, "pol_bind: compact layout needs 4 <= stored_cells < 2^31");
  if (b->layout_flags & ~POL_LAYOUT_NO_X_FACE)
    return fail(c, POL_ERR_INVALID, "pol_bind: unknown layout_flags bits");
  if (b->chunk_records && ((uintptr_t)b->chunk_records & 15))
    return fail(c, POL_ERR_INVALID, "pol_bind: chunk_records must be 16-byte aligned");
  const Grid wg = listed ? make_grid((uint64_t)b->n_chunks << cshift) : c->grid;
  if (wg.nblk > c->part_cap) return fail(c, POL_ERR_INVALID, "pol_bind: work grid exceeds the reduction scratch");
  c->b = *b;
  c->b.chunk_cells = cc;
  c->cshift = cshift;
  c->P = b->chunk_records ? b->stored_cells : c->N;
  if (!(o & POL_USE_FLAGS)) c->b.cell_flags = nullptr;
  c->wgrid = wg;
  c->bound = true;
  c->cur = 0;
  c->err.clear();
  return POL_OK;
}

int pol_bind(pol_ctx* c, const pol_buffers* b) {
  if (!c || !b) return POL_ERR_INVALID;
  return bind_buffers(c, b, true);
}

int pol_set_params(pol_ctx* c, const pol_desc* d) {
  if (!c || !d) return POL_ERR_INVALID;
  if (d->nx != c->d.nx || d->ny != c->d.ny || d->nz != c->d.nz || d->storage != c->d.storage ||
      (d->options & ~POL_LES) != (c->d.options & ~POL_LES) || d->device != c->d.device)
    return fail(c, POL_ERR_INVALID,
                "pol_set_params: geometry, storage, options (all but POL_LES) and device are fixed per context");
  if ((d->options & POL_LES) && (!(d->cs_smag >= 0.0) || !std::isfinite(d->cs_smag) || !(d->nu_sgs_max >= 0.0)))
    return fail(c, POL_ERR_INVALID, "pol_set_params: LES needs finite cs_smag >= 0 and nu_sgs_max >= 0");
  if (!(d->tau > 0.5) || !std::isfinite(d->tau)) return fail(c, POL_ERR_INVALID, "pol_set_params: tau must be > 1/2");
  c->d = *d;
  return POL_OK;
}

int pol_current(const pol_ctx* c, int* index) {
  if (!c || !index) return POL_ERR_INVALID;
  *index = c->cur;
  return POL_OK;
}

int pol_set_current(pol_ctx* c, int index) {
  if (!c || (index != 0 && index != 1)) return POL_ERR_INVALID;
  c->cur = index;
  return POL_OK;
}

static StepArgs make_args(const pol_ctx* c, const void* fin, void* fout, uint32_t what);

int pol_init_equilibrium(pol_ctx* c, const void* rho, const void* u, void* stream) {
  if (!c) return POL_ERR_INVALID;
  if (!c->bound) return fail(c, POL_ERR_STATE, "pol_init_equilibrium: buffers not bound");
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) {
    const StepArgs a = make_args(c, nullptr, c->b.f[c->cur], 0);
    e = launch_init_equilibrium(c->d.storage, a, rho, u, c->wgrid, (hipStream_t)stream);
  }
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_init_equilibrium");
}

static StepArgs make_args(const pol_ctx* c, const void* fin, void* fout, uint32_t what) {
  StepArgs a{};
  const pol_desc& d = c->d;
  a.fin = fin;
  a.fout = fout;
  a.flags = (d.options & POL_USE_FLAGS) ? c->b.cell_flags : nullptr;
  a.force = c->b.force;
  a.perm = c->b.perm;
  a.pdrag = c->b.particle_drag;
  a.nu_sgs = c->b.nu_sgs;
  a.rho = c->b.rho;
  a.u = c->b.u;
  a.partial = c->partial;
  a.N = c->N;
  a.nx = d.nx;
  a.ny = d.ny;
  a.nz = d.nz;
  a.nxy = (uint32_t)d.nx * d.ny;
  a.dnx = make_fastdiv(d.nx);
  a.dny = make_fastdiv(d.ny);
  a.chunks = c->b.active_chunks;
  a.records = c->b.chunk_records;
  a.P = c->P;
  a.n_items = c->b.active_chunks ? c->b.n_chunks << c->cshift : c->N;
  a.cshift = c->cshift;
  a.xface = (c->b.layout_flags & POL_LAYOUT_NO_X_FACE) ? 0u : 1u;
  // Streaming (non-temporal) stores once a step's population traffic (read +
  // write) exceeds the 256 MiB last-level cache: at 224^3 they cut the V60
  // step by 7 % (FP32 compact), 13 % (FP16) and 8 % (FP64); at 128^3, where
  // the step still caches, they cost 25 %.
  a.stream_store = 2ull * a.n_items * Q * storage_bytes(d.storage) > (256ull << 20) ? 1u : 0u;
  // Streaming steps sweep one memory front: linear blocks over the dense
  // grid, groups of 4 consecutive blocks per XCD over the compact chunk list.
  // A step that stays in the last-level cache keeps each XCD on one
  // contiguous range, whose neighbour rows then hit its own L2 (128^3 V60
  // FP32: 0.0203 ms against 0.0223 ms linear).
  a.order = a.stream_store ? (a.records ? 2u : ORDER_LINEAR) : ORDER_XCD_RANGES;
  a.per_xcd = c->wgrid.per_xcd;
  a.nblk = c->wgrid.nblk;
  a.options = d.options;
  a.what = what;
  a.tau0 = d.tau;
  a.nu0 = (d.tau - 0.5) / 3.0;
  a.les_c = 18.0 * std::sqrt(2.0) * d.cs_smag * d.cs_smag;
  a.tau_max = d.tau + 3.0 * d.nu_sgs_max;
  for (int i = 0; i < 3; ++i) a.G[i] = (d.options & POL_GRAVITY) ? d.gravity[i] : 0.0;
  a.K = d.darcy_K;
  a.CF = d.forch_CF;
  a.inlet_rho = d.inlet_rho;
  for (int i = 0; i < 3; ++i) a.inlet_u[i] = d.inlet_u[i];
  a.outlet_rho = d.outlet_rho;
  a.pf.tau0 = (float)a.tau0;
  a.pf.omega0 = (float)(1.0 / a.tau0);
  a.pf.nu0 = (float)a.nu0;
  a.pf.les_c = (float)a.les_c;
  a.pf.tau_max = (float)a.tau_max;
  for (int i = 0; i < 3; ++i) a.pf.G[i] = (float)a.G[i];
  a.pf.K = (float)a.K;
  a.pf.CF = (float)a.CF;
  a.pf.inlet_rho = (float)a.inlet_rho;
  for (int i = 0; i < 3; ++i) a.pf.inlet_u[i] = (float)a.inlet_u[i];
  a.pf.outlet_rho = (float)a.outlet_rho;
  return a;
}

static int check_what(pol_ctx* c, uint32_t what, double* red_out, const char* fn) {
  if ((what & POL_WANT_MACRO) && (!c->b.rho || !c->b.u))
    return fail(c, POL_ERR_STATE, std::string(fn) + ": macro output requested but rho/u not bound");
  if ((what & POL_WANT_REDUCE) && !red_out) return fail(c, POL_ERR_INVALID, std::string(fn) + ": red_out is NULL");
  return POL_OK;
}

static int finish_reduce(pol_ctx* c, double* red_out, hipStream_t s, const char* fn) {
  hipError_t e = launch_reduce_final(c->partial, c->wgrid.nblk, c->red_dev, s);
  if (e == hipSuccess) e = hipMemcpyAsync(red_out, c->red_dev, POL_NRED * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, fn);
}

// The reference's finite-difference eddy viscosity from the velocity u
// (real [3][cells]) into the bound nu_sgs buffer.
static hipError_t les_fd_pass(const pol_ctx* c, const void* u, hipStream_t s) {
  const pol_desc& d = c->d;
  LesFdArgs a{};
  a.u = u;
  a.flags = (d.options & POL_USE_FLAGS) ? c->b.cell_flags : nullptr;
  a.phase = c->b.les_phase;
  a.nu = c->b.nu_sgs;
  a.nx = d.nx;
  a.ny = d.ny;
  a.nz = d.nz;
  a.nxy = (uint32_t)d.nx * d.ny;
  a.N = c->N;
  a.dnx = make_fastdiv(d.nx);
  a.dny = make_fastdiv(d.ny);
  a.cs2 = (float)(d.cs_smag * d.cs_smag);  // (C_s * filter width 1)^2
  a.nu_max = (float)d.nu_sgs_max;
  return launch_les_fd(d.storage == POL_F64, a, s);
}

static bool les_fd(const pol_ctx* c) { return (c->d.options & (POL_LES | POL_LES_FD)) == (POL_LES | POL_LES_FD); }

int pol_les_viscosity(pol_ctx* c, const void* u_field, void* stream) {
  if (!c) return POL_ERR_INVALID;
  if (!c->bound) return fail(c, POL_ERR_STATE, "pol_les_viscosity: buffers not bound");
  if (!les_fd(c)) return fail(c, POL_ERR_STATE, "pol_les_viscosity: needs POL_LES | POL_LES_FD");
  const void* u = u_field ? u_field : c->b.u;
  if (!u) return fail(c, POL_ERR_INVALID, "pol_les_viscosity: no velocity field");
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) e = les_fd_pass(c, u, (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_les_viscosity");
}

int pol_step_ex(pol_ctx* c, int nsteps, uint32_t what, double* red_out, void* stream) {
  if (!c || nsteps < 0) return POL_ERR_INVALID;
  if (!c->bound) return fail(c, POL_ERR_STATE, "pol_step: buffers not bound");
  const bool fd = les_fd(c);
  int st = check_what(c, what | (fd ? POL_WANT_MACRO : 0u), red_out, "pol_step");
  if (st) return st;
  if (nsteps == 0) return POL_OK;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSetDevice(c->d.device);
  for (int i = 0; i < nsteps && e == hipSuccess; ++i) {
    // finite-difference LES: nu_sgs from the u the previous step wrote, and
    // this step writes u for the next (legacy/lbm_solver.py:856-861)
    if (fd) e = les_fd_pass(c, c->b.u, s);
    if (e != hipSuccess) break;
    const uint32_t w = ((i == nsteps - 1) ? what : 0u) | (fd ? POL_WANT_MACRO : 0u);
    StepArgs a = make_args(c, c->b.f[c->cur], c->b.f[1 - c->cur], w);
    e = launch_stream_collide(c->d.storage, a, true, c->wgrid, s);
    if (e == hipSuccess) c->cur = 1 - c->cur;
  }
  if (e != hipSuccess) return hip_fail(c, e, "pol_step");
  if (what & POL_WANT_REDUCE) return finish_reduce(c, red_out, s, "pol_step");
  return POL_OK;
}

int pol_step(pol_ctx* c, int nsteps, void* stream) { return pol_step_ex(c, nsteps, 0u, nullptr, stream); }

int pol_stream_only(pol_ctx* c, int nsteps, void* stream) {
  if (!c || nsteps < 0) return POL_ERR_INVALID;
  if (!c->bound) return fail(c, POL_ERR_STATE, "pol_stream_only: buffers not bound");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSetDevice(c->d.device);
  for (int i = 0; i < nsteps && e == hipSuccess; ++i) {
    const StepArgs a = make_args(c, c->b.f[c->cur], c->b.f[1 - c->cur], 0u);
    e = launch_stream_only(c->d.storage, a, c->wgrid, s);
    if (e == hipSuccess) c->cur = 1 - c->cur;
  }
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_stream_only");
}

int pol_step_from(pol_ctx* c, const void* fin, void* fout, uint32_t what, double* red_out, void* stream) {
  if (!c || !fin || !fout || fin == fout) return POL_ERR_INVALID;
  if (!c->bound) return fail(c, POL_ERR_STATE, "pol_step_from: buffers not bound");
  if (les_fd(c)) return fail(c, POL_ERR_UNSUPPORTED, "pol_step_from: not with the finite-difference LES");
  int st = check_what(c, what, red_out, "pol_step_from");
  if (st) return st;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSetDevice(c->d.device);
  StepArgs a = make_args(c, fin, fout, what);
  if (e == hipSuccess) e = launch_stream_collide(c->d.storage, a, true, c->wgrid, s);
  if (e != hipSuccess) return hip_fail(c, e, "pol_step_from");
  if (what & POL_WANT_REDUCE) return finish_reduce(c, red_out, s, "pol_step_from");
  return POL_OK;
}

int pol_macroscopic(pol_ctx* c, uint32_t what, double* red_out, void* stream) {
  if (!c) return POL_ERR_INVALID;
  if (!c->bound) return fail(c, POL_ERR_STATE, "pol_macroscopic: buffers not bound");
  int st = check_what(c, what, red_out, "pol_macroscopic");
  if (st) return st;
  if (!what) return POL_OK;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSetDevice(c->d.device);
  StepArgs a = make_args(c, c->b.f[c->cur], nullptr, what);
  if (e == hipSuccess) e = launch_stream_collide(c->d.storage, a, false, c->wgrid, s);
  if (e != hipSuccess) return hip_fail(c, e, "pol_macroscopic");
  if (what & POL_WANT_REDUCE) return finish_reduce(c, red_out, s, "pol_macroscopic");
  return POL_OK;
}

int pol_reduce(pol_ctx* c, double* red_out, void* stream) {
  return pol_macroscopic(c, POL_WANT_REDUCE, red_out, stream);
}

// ---- allocation-free reduction --------------------------------------------
// Per-block FP64 partials of pol_reduce_noalloc.  Each GPU the library runs on
// holds its own copy of this array until the process ends.  Host code never
// passes the array's name to a kernel (on the host it is a stand-in, not a
// device address): it uses the per-device address hipGetSymbolAddress
// returns, looked up once per device with that device current.
static __device__ __align__(256) double scratch[(size_t)POL_NOALLOC_MAX_BLOCKS * POL_NRED];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_dev[kMaxDevices] = {};

static int free_fail(int code, const std::string& msg) {
  g_free_err = msg;
  return code;
}

// The device address of `scratch` on the current device `dev`.
static hipError_t scratch_address(int dev, double** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[dev] = (double*)p;
  }
  *out = g_scratch_dev[dev];
  return hipSuccess;
}

int pol_reduce_noalloc(const pol_desc* desc, const pol_buffers* bufs, double* red_out_dev, void* stream) {
  if (!desc || !bufs || !red_out_dev) return free_fail(POL_ERR_INVALID, "pol_reduce_noalloc: null argument");
  const pol_desc& d = *desc;
  if (!desc_ok(d)) return free_fail(POL_ERR_INVALID, "pol_reduce_noalloc: invalid descriptor");
  // the context on the stack holds descriptors only (its device pointers stay null)
  pol_ctx c;
  c.d = d;
  const uint64_t N = (uint64_t)d.nx * d.ny * d.nz;
  c.N = (uint32_t)N;
  c.grid = make_grid(N);
  c.wgrid = c.grid;
  c.part_cap = POL_NOALLOC_MAX_BLOCKS;
  const bool listed = bufs->active_chunks != nullptr;
  const uint32_t cc = bufs->chunk_cells ? bufs->chunk_cells : POL_CHUNK;
  const uint64_t items = listed ? (uint64_t)bufs->n_chunks * cc : N;
  if ((items + BLOCK - 1) / BLOCK > POL_NOALLOC_MAX_BLOCKS)
    return free_fail(POL_ERR_INVALID, "pol_reduce_noalloc: " + std::to_string(items) +
                                          " work items exceed the static scratch (at most " +
                                          std::to_string((uint64_t)POL_NOALLOC_MAX_BLOCKS * BLOCK) + ")");
  int st = bind_buffers(&c, bufs, false);
  if (st) return free_fail(st, "pol_reduce_noalloc: " + c.err);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return free_fail(POL_ERR_NO_DEVICE, "pol_reduce_noalloc: no device");
  if (d.device < 0 || d.device >= ndev || d.device >= kMaxDevices)
    return free_fail(POL_ERR_INVALID, "pol_reduce_noalloc: device index out of range");
  hipStream_t s = (hipStream_t)stream;
  double* part = nullptr;
  hipError_t e = hipSetDevice(d.device);
  if (e == hipSuccess) e = scratch_address(d.device, &part);
  if (e == hipSuccess) {
    StepArgs a = make_args(&c, c.b.f[0], nullptr, POL_WANT_REDUCE);
    a.partial = part;  // every block writes all of its POL_NRED partials
    e = launch_stream_collide(d.storage, a, false, c.wgrid, s);
  }
  if (e == hipSuccess) e = launch_reduce_final(part, c.wgrid.nblk, red_out_dev, s);
  if (e != hipSuccess)
    return free_fail(POL_ERR_HIP, std::string("pol_reduce_noalloc: ") + hipGetErrorName(e) + " (" +
                                      hipGetErrorString(e) + ")");
  return POL_OK;
}

// ---- particles ----------------------------------------------------------
int pol_particles_bind(pol_ctx* c, const pol_particle_desc* d, const pol_particle_buffers* b) {
  if (!c || !d || !b) return POL_ERR_INVALID;
  if (d->count < 0) return fail(c, POL_ERR_INVALID, "pol_particles_bind: negative count");
  if (d->count > (1 << 27)) return fail(c, POL_ERR_INVALID, "pol_particles_bind: at most 2^27 particles (8 lanes each)");
  if (d->reaction != POL_REACTION_IMPLICIT && d->reaction != POL_REACTION_EXPLICIT)
    return fail(c, POL_ERR_INVALID, "pol_particles_bind: unknown reaction mode");
  const bool expl = d->reaction == POL_REACTION_EXPLICIT;
  if (!c->bound) return fail(c, POL_ERR_STATE, "pol_particles_bind: bind forward buffers first");
  if (!expl && !(c->d.options & POL_PARTICLE_FIELD))
    return fail(c, POL_ERR_STATE, "pol_particles_bind: the implicit reaction needs POL_PARTICLE_FIELD buffers");
  if (expl && !b->reaction_field)
    return fail(c, POL_ERR_INVALID, "pol_particles_bind: the explicit reaction needs reaction_field");
  if (d->count > 0 && (!b->pos || !b->vel || !b->radius || !b->mass || !b->active || !b->drag_new || !b->drag_old ||
                       !b->drag || !b->u_fluid || !b->re_p || !b->cd || (!expl && !b->drag_field) || !b->u_field))
    return fail(c, POL_ERR_INVALID, "pol_particles_bind: missing particle buffer");
  if (!(d->relax_alpha >= 0.0 && d->relax_alpha <= 1.0)) return fail(c, POL_ERR_INVALID, "pol_particles_bind: alpha outside [0,1]");
  if (!(d->dt > 0.0) || !(d->max_disp > 0.0)) return fail(c, POL_ERR_INVALID, "pol_particles_bind: dt and max_disp must be > 0");
  if (d->drag_law != POL_DRAG_TEXTBOOK && d->drag_law != POL_DRAG_REFERENCE)
    return fail(c, POL_ERR_INVALID, "pol_particles_bind: unknown drag_law");
  // A coupling footprint still in the drag field belongs to the old binding:
  // it stays tracked when the positions and the field are the same buffers,
  // else it is cleared now (ordered after the work already queued).
  bool keep_coupled = false;
  if (c->particles && c->pcoupled) {
    if (c->pa.pos == b->pos && c->pa.drag_field == (expl ? nullptr : b->drag_field) && c->pa.count == d->count) {
      keep_coupled = true;
    } else {
      hipError_t e = hipSetDevice(c->d.device);
      if (e == hipSuccess) e = launch_particles_clear_footprint(c->pa, nullptr);
      if (e == hipSuccess) e = hipDeviceSynchronize();
      if (e != hipSuccess) return hip_fail(c, e, "pol_particles_bind");
    }
  }
  ParticleArgs& p = c->pa;
  p = ParticleArgs{};
  p.count = d->count;
  p.real_is_double = c->d.storage == POL_F64;
  p.drag_law = d->drag_law;
  p.nx = c->d.nx;
  p.ny = c->d.ny;
  p.nz = c->d.nz;
  p.nxy = (uint32_t)c->d.nx * c->d.ny;
  p.N = c->N;
  p.flags = (c->d.options & POL_USE_FLAGS) ? c->b.cell_flags : nullptr;
  p.pos = b->pos;
  p.vel = b->vel;
  p.radius = b->radius;
  p.mass = b->mass;
  p.active = (int32_t*)b->active;
  p.drag_new = b->drag_new;
  p.drag_old = b->drag_old;
  p.drag = b->drag;
  p.u_fluid = b->u_fluid;
  p.re_p = b->re_p;
  p.cd = b->cd;
  p.drag_field = expl ? nullptr : b->drag_field;
  p.reaction = expl ? b->reaction_field : nullptr;
  p.u_field = b->u_field;
  p.rho_fluid = d->rho_fluid;
  p.nu_fluid = d->nu_fluid;
  p.drag_max_per_mass = d->drag_max_per_mass;
  p.alpha = d->relax_alpha;
  for (int i = 0; i < 3; ++i) p.g[i] = d->gravity[i];
  p.dt = d->dt;
  p.max_disp = d->max_disp;
  p.wall_damping = d->wall_damping;
  c->particles = true;
  c->pcoupled = keep_coupled;
  return POL_OK;
}

// One coupling pass: the implicit form first clears the footprint of a
// coupling not yet consumed by an advance (so a repeated couple recomputes
// instead of adding twice); the explicit form clears its reaction field, as
// compute_two_way_coupling_forces does (coffee_particles.py:1109-1111).
static hipError_t couple_pass(pol_ctx* c, const ParticleArgs& pa, bool relax, hipStream_t s) {
  hipError_t e = hipSuccess;
  if (pa.reaction) {
    e = launch_zero_real(pa.real_is_double, pa.reaction, 3 * (size_t)c->N, s);
  } else if (c->pcoupled) {
    e = launch_particles_clear_footprint(pa, s);
  }
  if (e == hipSuccess) e = launch_particles_couple(pa, relax, s);
  if (e == hipSuccess && !pa.reaction) c->pcoupled = true;
  return e;
}

int pol_particles_couple(pol_ctx* c, void* stream) {
  if (!c) return POL_ERR_INVALID;
  if (!c->particles) return fail(c, POL_ERR_STATE, "pol_particles_couple: no particles bound");
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) e = couple_pass(c, c->pa, true, (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_particles_couple");
}

int pol_particles_compute_coupling(pol_ctx* c, const void* u_field, void* stream) {
  if (!c) return POL_ERR_INVALID;
  if (!c->particles) return fail(c, POL_ERR_STATE, "pol_particles_compute_coupling: no particles bound");
  ParticleArgs pa = c->pa;
  if (u_field) pa.u_field = u_field;
  if (pa.count > 0 && !pa.u_field) return fail(c, POL_ERR_INVALID, "pol_particles_compute_coupling: no velocity field");
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) e = couple_pass(c, pa, false, (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_particles_compute_coupling");
}

int pol_particles_relax(pol_ctx* c, double alpha, void* stream) {
  if (!c) return POL_ERR_INVALID;
  if (!c->particles) return fail(c, POL_ERR_STATE, "pol_particles_relax: no particles bound");
  if (!(alpha >= 0.0 && alpha <= 1.0)) return fail(c, POL_ERR_INVALID, "pol_particles_relax: alpha outside [0,1]");
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) e = launch_particles_relax(c->pa, alpha, (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_particles_relax");
}

int pol_particles_clear_reaction(pol_ctx* c, void* stream) {
  if (!c) return POL_ERR_INVALID;
  if (!c->particles) return fail(c, POL_ERR_STATE, "pol_particles_clear_reaction: no particles bound");
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) {
    if (c->pa.reaction) {
      e = launch_zero_real(c->pa.real_is_double, c->pa.reaction, 3 * (size_t)c->N, (hipStream_t)stream);
    } else if (c->pcoupled) {
      e = launch_particles_clear_footprint(c->pa, (hipStream_t)stream);
      if (e == hipSuccess) c->pcoupled = false;
    }
  }
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_particles_clear_reaction");
}

int pol_add_reaction_forces(pol_ctx* c, const void* reaction_field, void* stream) {
  if (!c) return POL_ERR_INVALID;
  if (!c->bound) return fail(c, POL_ERR_STATE, "pol_add_reaction_forces: buffers not bound");
  if (!(c->d.options & POL_FORCE_FIELD) || !c->b.force)
    return fail(c, POL_ERR_STATE, "pol_add_reaction_forces: the lattice has no body-force field (POL_FORCE_FIELD)");
  const void* r = reaction_field ? reaction_field : (c->particles ? c->pa.reaction : nullptr);
  if (!r) return fail(c, POL_ERR_INVALID, "pol_add_reaction_forces: no reaction field");
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess)
    e = launch_add_field_fluid(c->d.storage == POL_F64, c->b.force, r, c->b.cell_flags, c->N, (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_add_reaction_forces");
}

int pol_particles_advance(pol_ctx* c, void* stream) {
  if (!c) return POL_ERR_INVALID;
  if (!c->particles) return fail(c, POL_ERR_STATE, "pol_particles_advance: no particles bound");
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) e = launch_particles_advance(c->pa, (hipStream_t)stream);
  if (e == hipSuccess) c->pcoupled = false;
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_particles_advance");
}

int pol_coupled_step(pol_ctx* c, int nsteps, uint32_t what, double* red_out, void* stream) {
  if (!c || nsteps < 0) return POL_ERR_INVALID;
  if (!c->bound) return fail(c, POL_ERR_STATE, "pol_coupled_step: buffers not bound");
  if (!c->particles) return fail(c, POL_ERR_STATE, "pol_coupled_step: no particles bound");
  if (!c->b.rho || !c->b.u) return fail(c, POL_ERR_INVALID, "pol_coupled_step: the coupling needs the rho/u buffers");
  const bool expl = c->pa.reaction != nullptr;
  if (expl && !((c->d.options & POL_FORCE_FIELD) && c->b.force))
    return fail(c, POL_ERR_STATE, "pol_coupled_step: the explicit reaction needs the body-force field (POL_FORCE_FIELD)");
  int st = check_what(c, what, red_out, "pol_coupled_step");
  if (st) return st;
  if (nsteps == 0) return POL_OK;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSetDevice(c->d.device);
  const bool dbl = c->d.storage == POL_F64;
  const bool fd = les_fd(c);
  for (int i = 0; i < nsteps && e == hipSuccess; ++i) {
    e = couple_pass(c, c->pa, true, s);
    if (e == hipSuccess && fd) e = les_fd_pass(c, c->b.u, s);
    if (e == hipSuccess && expl) {
      // legacy/lbm_solver.py:1488,1499: clear_body_force, then add the reaction
      e = launch_zero_real(dbl, c->b.force, 3 * (size_t)c->N, s);
      if (e == hipSuccess) e = launch_add_field_fluid(dbl, c->b.force, c->pa.reaction, c->b.cell_flags, c->N, s);
    }
    if (e != hipSuccess) break;
    const uint32_t w = POL_WANT_MACRO | ((i == nsteps - 1) ? what : 0u);
    StepArgs a = make_args(c, c->b.f[c->cur], c->b.f[1 - c->cur], w);
    e = launch_stream_collide(c->d.storage, a, true, c->wgrid, s);
    if (e != hipSuccess) break;
    c->cur = 1 - c->cur;
    e = launch_particles_advance(c->pa, s);
    if (e == hipSuccess) c->pcoupled = false;
  }
  if (e != hipSuccess) return hip_fail(c, e, "pol_coupled_step");
  if (what & POL_WANT_REDUCE) return finish_reduce(c, red_out, s, "pol_coupled_step");
  return POL_OK;
}

int pol_particle_index(const float* pos, int32_t n, int32_t nx, int32_t ny, int32_t nz, int32_t* idx, float* frac,
                       void* stream) {
  if (n < 0 || nx < 2 || ny < 2 || nz < 2) return POL_ERR_INVALID;
  if (n > 0 && (!pos || !idx || !frac)) return POL_ERR_INVALID;
  hipError_t e = launch_particle_index(pos, n, nx, ny, nz, idx, frac, (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : POL_ERR_HIP;
}

// ---- adjoint ---------------------------------------------------------------
int pol_adjoint_bind(pol_ctx* c, const pol_adjoint_buffers* b) {
  if (!c || !b) return POL_ERR_INVALID;
  if (c->d.storage != POL_F32 && c->d.storage != POL_F64)
    return fail(c, POL_ERR_UNSUPPORTED, "pol_adjoint_bind: adjoint supports F32 and F64 storage");
  if (c->d.options & POL_PARTICLE_FIELD)
    return fail(c, POL_ERR_UNSUPPORTED, "pol_adjoint_bind: the adjoint does not cover the particle-drag field");
  if (c->d.options & POL_LES_FD)
    return fail(c, POL_ERR_UNSUPPORTED, "pol_adjoint_bind: the adjoint covers the local-stress LES only");
  if (!(c->d.options & POL_USE_FLAGS) || !c->bound)
    return fail(c, POL_ERR_STATE, "pol_adjoint_bind: bind forward buffers with cell flags first");
  if (!b->lam[0] || !b->lam[1] || b->lam[0] == b->lam[1] || !b->rev_flags)
    return fail(c, POL_ERR_INVALID, "pol_adjoint_bind: two adjoint buffers and rev_flags required");
  c->ab = *b;
  c->acur = 0;
  c->adjoint = true;
  return POL_OK;
}

int pol_adjoint_reset(pol_ctx* c, void* stream) {
  if (!c) return POL_ERR_INVALID;
  if (!c->adjoint) return fail(c, POL_ERR_STATE, "pol_adjoint_reset: adjoint not bound");
  hipStream_t s = (hipStream_t)stream;
  const size_t rb = c->d.storage == POL_F64 ? 8 : 4;
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) e = hipMemsetAsync(c->ab.lam[0], 0, (size_t)POL_Q * c->P * rb, s);
  if (e == hipSuccess) e = hipMemsetAsync(c->ab.lam[1], 0, (size_t)POL_Q * c->P * rb, s);
  if (e == hipSuccess && c->ab.grad_perm) e = hipMemsetAsync(c->ab.grad_perm, 0, (size_t)c->N * sizeof(double), s);
  if (e == hipSuccess) e = hipMemsetAsync(c->adj_partial, 0, (size_t)c->wgrid.nblk * 4 * sizeof(double), s);
  if (e == hipSuccess) e = hipMemsetAsync(c->adj_acc, 0, 4 * sizeof(double), s);
  c->acur = 0;
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_adjoint_reset");
}

// One reverse step on stream s (the device is current).
static hipError_t adjoint_launch(pol_ctx* c, const void* f_state, int objective, hipStream_t s) {
  AdjointArgs A{};
  A.fw = make_args(c, f_state, nullptr, 0u);
  A.lam_in = c->ab.lam[c->acur];
  A.lam_out = c->ab.lam[1 - c->acur];
  A.grad_perm = (c->d.options & POL_PERM_FIELD) ? c->ab.grad_perm : nullptr;
  A.partial = c->adj_partial;
  A.objective = objective;
  const hipError_t e = launch_adjoint(c->d.storage, A, c->ab.rev_flags, c->wgrid, s);
  if (e == hipSuccess) c->acur = 1 - c->acur;
  return e;
}

int pol_adjoint_step(pol_ctx* c, const void* f_state, int objective, void* stream) {
  if (!c || !f_state) return POL_ERR_INVALID;
  if (!c->adjoint) return fail(c, POL_ERR_STATE, "pol_adjoint_step: adjoint not bound");
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) e = adjoint_launch(c, f_state, objective, (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_adjoint_step");
}

int pol_adjoint_sweep(pol_ctx* c, const void* states, uint64_t stride_bytes, int count, int objective_first,
                      void* stream) {
  if (!c || !states || count < 0) return POL_ERR_INVALID;
  if (!c->adjoint) return fail(c, POL_ERR_STATE, "pol_adjoint_sweep: adjoint not bound");
  if (count > 1 && stride_bytes == 0) return fail(c, POL_ERR_INVALID, "pol_adjoint_sweep: zero state stride");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSetDevice(c->d.device);
  for (int i = count - 1; i >= 0 && e == hipSuccess; --i)
    e = adjoint_launch(c, (const char*)states + (uint64_t)i * stride_bytes, i >= objective_first ? 1 : 0, s);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_adjoint_sweep");
}

int pol_step_chain(pol_ctx* c, void* states, uint64_t stride_bytes, int count, void* stream) {
  if (!c || !states || count < 0) return POL_ERR_INVALID;
  if (!c->bound) return fail(c, POL_ERR_STATE, "pol_step_chain: buffers not bound");
  if (les_fd(c)) return fail(c, POL_ERR_UNSUPPORTED, "pol_step_chain: not with the finite-difference LES");
  if (count > 1 && stride_bytes == 0) return fail(c, POL_ERR_INVALID, "pol_step_chain: zero state stride");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSetDevice(c->d.device);
  for (int i = 1; i < count && e == hipSuccess; ++i) {
    const StepArgs a = make_args(c, (const char*)states + (uint64_t)(i - 1) * stride_bytes,
                                 (char*)states + (uint64_t)i * stride_bytes, 0u);
    e = launch_stream_collide(c->d.storage, a, true, c->wgrid, s);
  }
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_step_chain");
}

int pol_adjoint_collect(pol_ctx* c, double* out4, void* stream) {
  if (!c || !out4) return POL_ERR_INVALID;
  if (!c->adjoint) return fail(c, POL_ERR_STATE, "pol_adjoint_collect: adjoint not bound");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) e = launch_adjoint_final(c->adj_partial, c->wgrid.nblk, c->adj_acc, s);
  if (e == hipSuccess) e = hipMemcpyAsync(out4, c->adj_acc, 4 * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_adjoint_collect");
}

// ---- field diagnostics --------------------------------------------------------
int pol_field_stats(pol_ctx* c, const pol_stats_opts* o, double* out, double* z_profile, void* stream) {
  if (!c || !o || !out) return POL_ERR_INVALID;
  if (!c->bound) return fail(c, POL_ERR_STATE, "pol_field_stats: buffers not bound");
  if (!c->b.rho || !c->b.u) return fail(c, POL_ERR_INVALID, "pol_field_stats: needs the rho / u buffers");
  hipStream_t st = (hipStream_t)stream;
  const pol_desc& d = c->d;
  StatsArgs s{};
  s.rho = c->b.rho;
  s.u = c->b.u;
  s.flags = (d.options & POL_USE_FLAGS) ? c->b.cell_flags : nullptr;
  s.prev_rho = o->prev_rho;
  s.prev_u = o->prev_u;
  s.phase = o->phase;
  s.nx = d.nx;
  s.ny = d.ny;
  s.nz = d.nz;
  s.nxy = (uint32_t)d.nx * d.ny;
  s.N = c->N;
  s.dnx = make_fastdiv(d.nx);
  s.tiles = c->stat_tiles;
  s.z_top = o->z_top;
  s.z_mid = o->z_mid;
  s.z_out = o->z_outlet;
  s.partial = c->stat_partial;
  s.zpart = c->stat_zpart;
  hipError_t e = hipSetDevice(d.device);
  if (e == hipSuccess) e = launch_field_stats(d.storage == POL_F64, s, c->stat_dev, c->stat_dev + POL_NSTAT, st);
  if (e == hipSuccess && o->with_populations) {
    const StepArgs a = make_args(c, c->b.f[c->cur], nullptr, 0u);
    e = launch_pop_nonfinite(d.storage, a, c->partial, c->stat_dev, c->wgrid, st);
  }
  // one transfer of statistics and profile into pinned memory, then host copies
  const size_t nstat = (size_t)POL_NSTAT + (z_profile ? d.nz : 0);
  if (e == hipSuccess) e = hipMemcpyAsync(c->stat_host, c->stat_dev, nstat * sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(c, e, "pol_field_stats");
  std::memcpy(out, c->stat_host, POL_NSTAT * sizeof(double));
  if (z_profile) std::memcpy(z_profile, c->stat_host + POL_NSTAT, (size_t)d.nz * sizeof(double));
  return POL_OK;
}

int pol_export_fields(pol_ctx* c, void* magnitude, void* pressure, void* stream) {
  if (!c || !magnitude || !pressure) return POL_ERR_INVALID;
  if (!c->bound) return fail(c, POL_ERR_STATE, "pol_export_fields: buffers not bound");
  if (!c->b.rho || !c->b.u) return fail(c, POL_ERR_INVALID, "pol_export_fields: needs the rho / u buffers");
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess)
    e = launch_export_fields(c->d.storage == POL_F64, c->b.rho, c->b.u, c->N, magnitude, pressure,
                             (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_export_fields");
}

int pol_stabilize(pol_ctx* c, double max_velocity, double min_density, double max_density, void* stream) {
  if (!c) return POL_ERR_INVALID;
  if (!c->bound) return fail(c, POL_ERR_STATE, "pol_stabilize: buffers not bound");
  if (!c->b.rho || !c->b.u) return fail(c, POL_ERR_INVALID, "pol_stabilize: needs the rho / u buffers");
  if (!(max_velocity > 0.0) || !(min_density > 0.0) || !(max_density >= min_density))
    return fail(c, POL_ERR_INVALID, "pol_stabilize: thresholds must satisfy 0 < min_density <= max_density, "
                                    "max_velocity > 0");
  const StepArgs a = make_args(c, c->b.f[c->cur], c->b.f[c->cur], 0u);
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess)
    e = launch_stabilize(c->d.storage, a, (float)max_velocity, (float)min_density, (float)max_density, c->wgrid,
                         (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_stabilize");
}

// ---- multiphase phase field -----------------------------------------------------
static MultiphaseArgs mp_args(const pol_ctx* c) {
  MultiphaseArgs a{};
  const pol_desc& d = c->d;
  a.phi = c->mb.phi[c->mcur];
  a.phi_out = c->mb.phi[1 - c->mcur];
  a.phi_new = c->mb.phi_new;
  a.mu = c->mb.mu;
  a.u = c->b.u;
  a.rho = c->b.rho;
  a.phase = c->mb.phase;
  a.force = (d.options & POL_FORCE_FIELD) ? c->b.force : nullptr;
  a.flags = (d.options & POL_USE_FLAGS) ? c->b.cell_flags : nullptr;
  a.nx = d.nx;
  a.ny = d.ny;
  a.nz = d.nz;
  a.nxy = (uint32_t)d.nx * d.ny;
  a.N = c->N;
  a.dnx = make_fastdiv(d.nx);
  a.dny = make_fastdiv(d.ny);
  const pol_multiphase_desc& m = c->md;
  a.sigma = (float)m.sigma;
  a.mobility = (float)m.mobility;
  a.dt = (float)m.dt;
  a.kappa = (float)(3.0 * m.sigma * m.interface_width / 8.0);
  a.rho_air = (float)m.rho_air;
  a.drho = (float)(m.rho_water - m.rho_air);
  return a;
}

int pol_multiphase_bind(pol_ctx* c, const pol_multiphase_desc* d, const pol_multiphase_buffers* b) {
  if (!c || !d || !b) return POL_ERR_INVALID;
  if (!c->bound || !c->b.rho || !c->b.u)
    return fail(c, POL_ERR_STATE, "pol_multiphase_bind: bind the context (rho and u buffers) first");
  if (!b->phi[0] || !b->phi[1] || b->phi[0] == b->phi[1] || !b->phi_new || !b->mu || !b->phase)
    return fail(c, POL_ERR_INVALID, "pol_multiphase_bind: two phi buffers, phi_new, mu and phase required");
  if (!std::isfinite(d->sigma) || !std::isfinite(d->mobility) || !std::isfinite(d->interface_width) ||
      !std::isfinite(d->rho_water) || !std::isfinite(d->rho_air) || !std::isfinite(d->dt))
    return fail(c, POL_ERR_INVALID, "pol_multiphase_bind: non-finite parameter");
  c->md = *d;
  c->mb = *b;
  c->mcur = 0;
  c->mp = true;
  return POL_OK;
}

int pol_multiphase_chemical_potential(pol_ctx* c, float* laplacian_phi, void* stream) {
  if (!c) return POL_ERR_INVALID;
  if (!c->mp) return fail(c, POL_ERR_STATE, "pol_multiphase_chemical_potential: phase field not bound");
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) e = launch_mp_chemical_potential(mp_args(c), laplacian_phi, (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_multiphase_chemical_potential");
}

int pol_multiphase_interface(pol_ctx* c, uint32_t flags, const pol_interface_fields* out, void* stream) {
  if (!c || (flags & ~POL_MP_APPLY)) return POL_ERR_INVALID;
  if (!c->mp) return fail(c, POL_ERR_STATE, "pol_multiphase_interface: phase field not bound");
  MultiphaseArgs a = mp_args(c);
  a.apply = (flags & POL_MP_APPLY) ? 1u : 0u;
  if (a.apply && !a.force)
    return fail(c, POL_ERR_STATE, "pol_multiphase_interface: POL_MP_APPLY needs the context's force field");
  bool any = false;
  if (out) {
    a.grad_phi = out->grad_phi;
    a.grad_mu = out->grad_mu;
    a.normal = out->normal;
    a.curvature = out->curvature;
    a.sforce = out->surface_force;
    any = a.grad_phi || a.grad_mu || a.normal || a.curvature || a.sforce;
  }
  if (!any && !a.apply) return POL_OK;
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) e = launch_mp_interface(c->d.storage == POL_F64, a, any, (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_multiphase_interface");
}

int pol_multiphase_step(pol_ctx* c, void* stream) {
  if (!c) return POL_ERR_INVALID;
  if (!c->mp) return fail(c, POL_ERR_STATE, "pol_multiphase_step: phase field not bound");
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) e = launch_mp_phase(c->d.storage == POL_F64, mp_args(c), (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(c, e, "pol_multiphase_step");
  c->mcur = 1 - c->mcur;
  return POL_OK;
}

int pol_multiphase_density(pol_ctx* c, void* stream) {
  if (!c) return POL_ERR_INVALID;
  if (!c->mp) return fail(c, POL_ERR_STATE, "pol_multiphase_density: phase field not bound");
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) e = launch_mp_density(c->d.storage == POL_F64, mp_args(c), (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_multiphase_density");
}

int pol_multiphase_current(pol_ctx* c, int* index) {
  if (!c || !index) return POL_ERR_INVALID;
  if (!c->mp) return fail(c, POL_ERR_STATE, "pol_multiphase_current: phase field not bound");
  *index = c->mcur;
  return POL_OK;
}

// ---- body-force producers ---------------------------------------------------------
static ForcingArgs forcing_args(const pol_ctx* c) {
  ForcingArgs f{};
  const pol_desc& d = c->d;
  f.force = (d.options & POL_FORCE_FIELD) ? c->b.force : nullptr;
  f.rho = c->b.rho;
  f.flags = (d.options & POL_USE_FLAGS) ? c->b.cell_flags : nullptr;
  f.nx = d.nx;
  f.ny = d.ny;
  f.nz = d.nz;
  f.nxy = (uint32_t)d.nx * d.ny;
  f.N = c->N;
  f.dnx = make_fastdiv(d.nx);
  f.dny = make_fastdiv(d.ny);
  return f;
}

static int pour_args(pol_ctx* c, const pol_pour_desc* d, PourArgs& p, const char* where) {
  if (!std::isfinite(d->center_x) || !std::isfinite(d->center_y) || !(d->radius >= 0.0) ||
      !std::isfinite(d->radius) || !std::isfinite(d->stream_height) || !std::isfinite(d->velocity) ||
      !std::isfinite(d->flow_rate) || !std::isfinite(d->dt))
    return fail(c, POL_ERR_INVALID, std::string(where) + ": non-finite pour parameter");
  p = PourArgs{};
  p.cx = (float)d->center_x;
  p.cy = (float)d->center_y;
  p.radius = (float)d->radius;
  p.zmin = (float)((double)d->height - d->stream_height);
  p.velocity = (float)d->velocity;
  p.flow_rate = (float)d->flow_rate;
  p.dt = (float)d->dt;
  p.height = d->height;
  // bounding box of |(i, j) - c| <= radius, zmin <= k <= height, one cell of
  // margin; bounds far outside the grid are clamped before the integer
  // conversion (a finite centre or radius of any size stays defined)
  auto lim = [](double v) { return std::min(std::max(v, -1.0e9), 1.0e9); };
  auto lo = [&](double v) { return (int64_t)std::floor(lim(v)) - 1; };
  auto hi = [&](double v) { return (int64_t)std::ceil(lim(v)) + 1; };
  const int64_t x0 = std::max<int64_t>(0, lo((double)p.cx - p.radius));
  const int64_t x1 = std::min<int64_t>(c->d.nx - 1, hi((double)p.cx + p.radius));
  const int64_t y0 = std::max<int64_t>(0, lo((double)p.cy - p.radius));
  const int64_t y1 = std::min<int64_t>(c->d.ny - 1, hi((double)p.cy + p.radius));
  const int64_t z0 = std::max<int64_t>(0, lo((double)p.zmin));
  const int64_t z1 = std::min<int64_t>(c->d.nz - 1, (int64_t)d->height);
  p.x0 = (int32_t)x0;
  p.y0 = (int32_t)y0;
  p.z0 = (int32_t)z0;
  p.bx = x1 >= x0 ? (int32_t)(x1 - x0 + 1) : 0;
  p.by = y1 >= y0 ? (int32_t)(y1 - y0 + 1) : 0;
  p.bz = z1 >= z0 ? (int32_t)(z1 - z0 + 1) : 0;
  if (p.bx == 0 || p.by == 0) p.bz = 0;
  return POL_OK;
}

int pol_pour_force(pol_ctx* c, const pol_pour_desc* d, void* stream) {
  if (!c || !d) return POL_ERR_INVALID;
  if (!c->bound) return fail(c, POL_ERR_STATE, "pol_pour_force: buffers not bound");
  const ForcingArgs f = forcing_args(c);
  if (!f.force) return fail(c, POL_ERR_STATE, "pol_pour_force: needs the context's force field (POL_FORCE_FIELD)");
  PourArgs p;
  int st = pour_args(c, d, p, "pol_pour_force");
  if (st) return st;
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) e = launch_pour(c->d.storage == POL_F64, f, p, nullptr, (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_pour_force");
}

int pol_pour_phase(pol_ctx* c, const pol_pour_desc* d, float* phi, void* stream) {
  if (!c || !d || !phi) return POL_ERR_INVALID;
  if (!c->bound) return fail(c, POL_ERR_STATE, "pol_pour_phase: buffers not bound");
  PourArgs p;
  int st = pour_args(c, d, p, "pol_pour_phase");
  if (st) return st;
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) e = launch_pour(c->d.storage == POL_F64, forcing_args(c), p, phi, (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_pour_phase");
}

int pol_density_drive(pol_ctx* c, const float* target, double rate, int half, double max_adjust, double lo, double hi,
                      void* stream) {
  if (!c || !target) return POL_ERR_INVALID;
  if (!c->bound || !c->b.rho) return fail(c, POL_ERR_STATE, "pol_density_drive: buffers not bound");
  if (!std::isfinite(rate) || !std::isfinite(max_adjust) || !(lo <= hi))
    return fail(c, POL_ERR_INVALID, "pol_density_drive: invalid parameters");
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess)
    e = launch_density_drive(c->d.storage == POL_F64, forcing_args(c), target, (float)rate, half != 0,
                             (float)max_adjust, (float)lo, (float)hi, (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_density_drive");
}

int pol_pressure_force(pol_ctx* c, float* pressure_force, int mode, double max_force, double cs2, void* stream) {
  if (!c || !pressure_force) return POL_ERR_INVALID;
  const bool st = (mode & POL_PRESSURE_SURFACE_TENSION) != 0;
  mode &= ~POL_PRESSURE_SURFACE_TENSION;
  if (mode < 0 || mode > 2) return fail(c, POL_ERR_INVALID, "pol_pressure_force: mode must be 0, 1 or 2");
  if (!c->bound || !c->b.rho) return fail(c, POL_ERR_STATE, "pol_pressure_force: buffers not bound");
  const ForcingArgs f = forcing_args(c);
  if (mode && !f.force)
    return fail(c, POL_ERR_STATE, "pol_pressure_force: modes 1 and 2 need the context's force field");
  if (st && mode == 0)
    return fail(c, POL_ERR_INVALID, "pol_pressure_force: the surface tension follows modes 1 and 2 only");
  if (st && !c->mp) return fail(c, POL_ERR_STATE, "pol_pressure_force: surface tension needs a bound phase field");
  MultiphaseArgs m{};
  if (st) {
    m = mp_args(c);
    m.apply = 1u;
  }
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess)
    e = launch_pressure_force(c->d.storage == POL_F64, f, st ? &m : nullptr, pressure_force, (float)max_force,
                              (float)cs2, mode, (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_pressure_force");
}

int pol_filter_blockage(pol_ctx* c, const uint32_t* cells, int32_t ncells, float* accumulated, float* blockage,
                        void* stream) {
  if (!c || ncells < 0 || (ncells > 0 && !cells) || !accumulated || !blockage) return POL_ERR_INVALID;
  if ((uint32_t)ncells > c->N)
    return fail(c, POL_ERR_INVALID, "pol_filter_blockage: more paper cells than grid cells");
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) e = launch_filter_blockage(cells, (uint32_t)ncells, accumulated, blockage, (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_filter_blockage");
}

int pol_filter_block_particles(pol_ctx* c, const uint8_t* zone, float* accumulated, const float* pos, float* vel,
                               const int32_t* active, int32_t count, double scale, uint32_t seed, void* stream) {
  if (!c || !zone || !accumulated || count < 0 || (count > 0 && (!pos || !vel || !active))) return POL_ERR_INVALID;
  if (!(scale > 0.0) || !std::isfinite(scale))
    return fail(c, POL_ERR_INVALID, "pol_filter_block_particles: scale must be positive");
  FilterParticleArgs a{};
  a.zone = zone;
  a.acc = accumulated;
  a.pos = pos;
  a.vel = vel;
  a.active = active;
  a.count = (uint32_t)count;
  a.nx = c->d.nx;
  a.ny = c->d.ny;
  a.nz = c->d.nz;
  a.seed = seed;
  a.scale = (float)scale;
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) e = launch_filter_block(a, (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_filter_block_particles");
}

// ---- thermal D3Q7 -----------------------------------------------------------
static ThermalArgs thermal_args(const pol_ctx* c) {
  ThermalArgs a{};
  const pol_desc& d = c->d;
  a.gin = c->tb.g[c->gcur];
  a.gout = c->tb.g[c->gcur];
  a.T = c->tb.temperature[c->tcur];
  a.Tout = c->tb.temperature[c->tcur];
  a.hs = c->tb.heat_source;
  a.hsb = (c->td.options & POL_THERMAL_RESET_SOURCE) ? c->tb.heat_source_base : nullptr;
  a.u = c->tb.velocity;
  a.k = c->tb.conductivity;
  a.nx = d.nx;
  a.ny = d.ny;
  a.nz = d.nz;
  a.nxy = (uint32_t)d.nx * d.ny;
  a.N = c->N;
  a.dnx = make_fastdiv(d.nx);
  a.dny = make_fastdiv(d.ny);
  a.omega_f = (float)c->td.omega;
  a.omega_d = c->td.omega;
  a.conv = (c->td.options & POL_THERMAL_CONVECTION) ? 1u : 0u;
  a.partial = c->partial;
  return a;
}

static int thermal_desc_ok(pol_ctx* c, const pol_thermal_desc* d, const char* where) {
  if (!(d->omega > 0.0 && d->omega < 2.0))
    return fail(c, POL_ERR_INVALID, std::string(where) + ": omega must lie in (0, 2)");
  if (d->options & ~(POL_THERMAL_CONVECTION | POL_THERMAL_RESET_SOURCE))
    return fail(c, POL_ERR_INVALID, std::string(where) + ": unknown options");
  return POL_OK;
}

int pol_thermal_bind(pol_ctx* c, const pol_thermal_desc* d, const pol_thermal_buffers* b) {
  if (!c || !d || !b) return POL_ERR_INVALID;
  int st = thermal_desc_ok(c, d, "pol_thermal_bind");
  if (st) return st;
  if (!b->g[0] || !b->g[1] || b->g[0] == b->g[1] || !b->temperature[0] || !b->temperature[1] ||
      b->temperature[0] == b->temperature[1] || !b->heat_source)
    return fail(c, POL_ERR_INVALID, "pol_thermal_bind: two g buffers, two temperature buffers and heat_source required");
  if ((d->options & POL_THERMAL_CONVECTION) && !b->velocity)
    return fail(c, POL_ERR_INVALID, "pol_thermal_bind: convection needs the velocity buffer");
  if ((d->options & POL_THERMAL_RESET_SOURCE) && !b->heat_source_base)
    return fail(c, POL_ERR_INVALID, "pol_thermal_bind: POL_THERMAL_RESET_SOURCE needs heat_source_base");
  c->td = *d;
  c->tb = *b;
  c->gcur = 0;
  c->tcur = 0;
  c->thermal = true;
  return POL_OK;
}

int pol_thermal_set(pol_ctx* c, const pol_thermal_desc* d) {
  if (!c || !d) return POL_ERR_INVALID;
  if (!c->thermal) return fail(c, POL_ERR_STATE, "pol_thermal_set: thermal buffers not bound");
  int st = thermal_desc_ok(c, d, "pol_thermal_set");
  if (st) return st;
  if ((d->options & POL_THERMAL_CONVECTION) && !c->tb.velocity)
    return fail(c, POL_ERR_INVALID, "pol_thermal_set: convection needs the velocity buffer");
  if ((d->options & POL_THERMAL_RESET_SOURCE) && !c->tb.heat_source_base)
    return fail(c, POL_ERR_INVALID, "pol_thermal_set: POL_THERMAL_RESET_SOURCE needs heat_source_base");
  c->td = *d;
  return POL_OK;
}

int pol_thermal_init(pol_ctx* c, void* stream) {
  if (!c) return POL_ERR_INVALID;
  if (!c->thermal) return fail(c, POL_ERR_STATE, "pol_thermal_init: thermal buffers not bound");
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) e = launch_thermal_init(c->d.storage == POL_F64, thermal_args(c), (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_thermal_init");
}

int pol_thermal_step(pol_ctx* c, int nsteps, int32_t* unstable_out, void* stream) {
  if (!c || nsteps < 0) return POL_ERR_INVALID;
  if (!c->thermal) return fail(c, POL_ERR_STATE, "pol_thermal_step: thermal buffers not bound");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess)
    e = launch_thermal_steps(c->d.storage == POL_F64, thermal_args(c), c->tb.g, c->tb.temperature, c->gcur, c->tcur,
                             nsteps, s);
  if (e == hipSuccess && unstable_out) {
    if (nsteps == 0) {
      *unstable_out = 0;
    } else {
      e = launch_thermal_fold(c->partial, (c->N + 255) / 256, 1, c->red_dev, s);
      double v = 0.0;
      if (e == hipSuccess) e = hipMemcpyAsync(&v, c->red_dev, sizeof(double), hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      *unstable_out = (int32_t)v;
    }
  }
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_thermal_step");
}

int pol_thermal_heat_flux(pol_ctx* c, void* flux, void* stream) {
  if (!c || !flux) return POL_ERR_INVALID;
  if (!c->thermal) return fail(c, POL_ERR_STATE, "pol_thermal_heat_flux: thermal buffers not bound");
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess) e = launch_thermal_flux(c->d.storage == POL_F64, thermal_args(c), flux, (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_thermal_heat_flux");
}

int pol_thermal_stats(pol_ctx* c, double* out3, void* stream) {
  if (!c || !out3) return POL_ERR_INVALID;
  if (!c->thermal) return fail(c, POL_ERR_STATE, "pol_thermal_stats: thermal buffers not bound");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess)
    e = launch_thermal_stats(c->d.storage == POL_F64, c->tb.temperature[c->tcur], c->N, c->partial, s);
  if (e == hipSuccess) e = launch_thermal_fold(c->partial, (c->N + 255) / 256, 3, c->red_dev, s);
  if (e == hipSuccess) e = hipMemcpyAsync(out3, c->red_dev, 3 * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_thermal_stats");
}

int pol_thermal_boundary(pol_ctx* c, int kind, const uint8_t* mask, double a, double b, void* stream) {
  if (!c || !mask || kind < 1 || kind > 3) return POL_ERR_INVALID;
  if (!c->thermal) return fail(c, POL_ERR_STATE, "pol_thermal_boundary: thermal buffers not bound");
  hipError_t e = hipSetDevice(c->d.device);
  if (e == hipSuccess)
    e = launch_thermal_boundary(c->d.storage == POL_F64, thermal_args(c), kind, mask, a, b, (hipStream_t)stream);
  return e == hipSuccess ? POL_OK : hip_fail(c, e, "pol_thermal_boundary");
}

}  // extern "C"
