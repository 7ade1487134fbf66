// Field diagnostics of the solver state as fused device reductions.
//
// Replaces (reference): the host-side numpy / Taichi reductions of
//   numerical_stability.py:53-110   (NaN / Inf counts, rho range, max |u|
//                                    over fluid cells, populations included)
//   numerical_stability.py:165-197  (emergency stabilisation)
//   lbm_diagnostics.py:151-460      (Mach / density quality, conservation,
//                                    regional |u|, outlet flow, z profile,
//                                    temporal change)
//   visualizer.py:120-167           (interior water / air mass, |u| range)
//   enhanced_visualizer.py:1487-1505 (|u| and pressure fields of the export)
// which copy rho, u (and f) to the host, about 180 MB per diagnostic step at
// 224^3.  Here one pass reads each field once on the device.
//
// Statistics kernel: a grid of (tiles, nz) blocks, block (t, z) looping over
// a strip of plane z; per-lane FP64 accumulators (integer counters for the
// counts, integer keys for the extremes of |u|) split between two wave
// pairs, wave64 shuffles, one partial per block per statistic, folded in
// block order by one block per statistic (deterministic).
// The per-plane density sums of the z profile fall out of the same partials.
// |u| is formed in FP32 as sqrt((ux*ux + uy*uy) + uz*uz) with explicit
// round-to-nearest operations (no contraction), the value the reference's
// float32 numpy expressions give per element.
#include "engine.h"
#include "cell_model.h"

namespace pol {

enum StatKind { SK_SUM = 0, SK_MIN = 1, SK_MAX = 2, SK_NANMAX = 3 };

__host__ __device__ constexpr int stat_kind(int s) {
  return (s == POL_STAT_MAX_ABS_DRHO || s == POL_STAT_MAX_UMAG)                                  ? SK_NANMAX
         : (s == POL_STAT_FLUID_RHO_MIN || s == POL_STAT_INT_UMIN)                               ? SK_MIN
         : (s == POL_STAT_FLUID_RHO_MAX || s == POL_STAT_FLUID_UMAX || s == POL_STAT_INT_UMAX) ? SK_MAX
                                                                                               : SK_SUM;
}
__host__ __device__ constexpr double stat_identity(int s) {
  return s == POL_STAT_FLUID_UMAX ? 0.0
         : stat_kind(s) == SK_MIN ? 1.0e300
         : (stat_kind(s) == SK_MAX || stat_kind(s) == SK_NANMAX) ? -1.0e300
                                                                 : 0.0;
}
__device__ inline double nanmax(double a, double b) { return (isnan(a) || isnan(b)) ? NAN : fmax(a, b); }
__device__ inline double stat_combine(int s, double a, double b) {
  switch (stat_kind(s)) {
    case SK_MIN: return fmin(a, b);
    case SK_MAX: return fmax(a, b);
    case SK_NANMAX: return nanmax(a, b);
    default: return a + b;
  }
}
template <int S>
__device__ inline double combine_c(double a, double b) {
  if constexpr (stat_kind(S) == SK_MIN) return fmin(a, b);
  else if constexpr (stat_kind(S) == SK_MAX) return fmax(a, b);
  else if constexpr (stat_kind(S) == SK_NANMAX) return nanmax(a, b);
  else return a + b;
}
__device__ inline double wave_combine(int s, double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = stat_combine(s, v, __shfl_xor(v, o, 64));
  return v;
}

// |u| in FP32 without contraction: ((ux*ux) + (uy*uy)) + (uz*uz), each
// operation rounded on its own (no FMA), then the correctly rounded sqrt
__device__ inline float umag_f32(float ux, float uy, float uz) {
#pragma clang fp contract(off)
  const float s = (ux * ux + uy * uy) + uz * uz;
  return __builtin_sqrtf(s);
}
__device__ inline double umag_f32(double ux, double uy, double uz) {
  return umag_f32((float)ux, (float)uy, (float)uz);
}

// Count statistics accumulate as integers (two counts in the 16-bit halves of
// one register per lane; reduced across the wave as 32-bit integers).
__host__ __device__ constexpr bool stat_is_count(int s) {
  return s == POL_STAT_CELLS || s == POL_STAT_ACTIVE || s == POL_STAT_MACH_GT || s == POL_STAT_FLUID ||
         s == POL_STAT_FLUID_RHO_NAN || s == POL_STAT_FLUID_RHO_INF || s == POL_STAT_FLUID_U_NAN ||
         s == POL_STAT_FLUID_U_INF || s == POL_STAT_INT_CELLS || s == POL_STAT_UMAG_ACTIVE;
}

// The statistics are split between the two wave pairs of a block, which walk
// the same cells of the block's strip of plane z: waves 0-1 (role A) keep
// the statistics of rho and of the velocity components (the sums over every
// cell and the non-solid density checks), waves 2-3 (role B) everything that
// needs |u| (its sums, Mach and activity, the non-solid and interior |u|
// statistics) and the temporal change.  Each lane then holds about half of
// the FP64 accumulators (twice the resident waves of a single set), |u| is
// formed once per cell, and each wave reduces only its own statistics; the
// second read of a cell comes from the cache.  The three regional |u| sums
// are the block's |u| sum placed by its (block-uniform) plane.  When a plane
// holds a multiple of four cells and a row at least four, lanes read four
// consecutive cells per load (16-byte loads of rho, u and the flags).
enum StatRole { ROLE_A = 0, ROLE_B = 1, ROLE_REGION = 2, ROLE_NONE = 3 };
__host__ __device__ constexpr int stat_role(int s) {
  return (s == POL_STAT_SUM_UMAG_TOP || s == POL_STAT_SUM_UMAG_MID || s == POL_STAT_SUM_UMAG_BOT) ? ROLE_REGION
         : (s == POL_STAT_FLUID_POP_NAN || s == POL_STAT_FLUID_POP_INF)                          ? ROLE_NONE
         : (s == POL_STAT_CELLS || s == POL_STAT_SUM_DRHO || s == POL_STAT_SUM_DRHO2 || s == POL_STAT_MAX_ABS_DRHO ||
            (s >= POL_STAT_MOM_X && s <= POL_STAT_MOM_Z) || (s >= POL_STAT_SUM_UX && s <= POL_STAT_SUM_UZ2) ||
            s == POL_STAT_OUTLET_DOWN || (s >= POL_STAT_FLUID && s <= POL_STAT_FLUID_RHO_MAX))
             ? ROLE_A
             : ROLE_B;
}

// Wave reduction of one statistic's lane values into sh[wid][S] (lane 0).
template <int S>
__device__ inline void wave_stat(double v, uint32_t cnt, double (*sh)[POL_NSTAT + 1], int wid, int lane) {
  if constexpr (stat_is_count(S)) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if (lane == 0) sh[wid][S] = (double)cnt;
  } else {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = combine_c<S>(v, __shfl_xor(v, o, 64));
    if (lane == 0) sh[wid][S] = v;
  }
}

// four consecutive values of a plane (p 16-byte aligned)
__device__ inline void load4(const float* p, float (&v)[4]) {
  const float4 t = *(const float4*)p;
  v[0] = t.x;
  v[1] = t.y;
  v[2] = t.z;
  v[3] = t.w;
}
__device__ inline void load4(const double* p, double (&v)[4]) {
  const double2 t0 = *(const double2*)p, t1 = *((const double2*)p + 1);
  v[0] = t0.x;
  v[1] = t0.y;
  v[2] = t1.x;
  v[3] = t1.y;
}
__device__ inline void load4(const uint32_t* p, uint32_t (&v)[4]) {
  const uint4 t = *(const uint4*)p;
  v[0] = t.x;
  v[1] = t.y;
  v[2] = t.z;
  v[3] = t.w;
}

// Mach = |u| / c_s in FP32, correctly rounded, by the method for a divisor
// known in advance (Brisebarre, Muller and Raina, IEEE Trans. Computers 53(8),
// 2004): q = fma(x, zh, x * zl) with zh = RN(1 / c_s), zl = RN(1 / c_s - zh).
// For c_s = float32(1 / sqrt(3)) this equals x / c_s for every x whose
// product x * zl is a normal number (x >= 2^-101, checked exhaustively over
// a binade on the host); below that both quotients are < 1e-33, under
// every Mach threshold.  x = inf would give inf - inf, so it keeps x.
__device__ inline float mach_of(float x) {
#pragma clang fp contract(off)
  constexpr float zh = 0x1.bb67bp+0f;      // RN(1 / c_s)
  constexpr float zl = -0x1.e9ecd6p-25f;   // RN(1 / c_s - zh)
  const float q = __builtin_fmaf(x, zh, x * zl);
  return __builtin_isinf(x) ? x : q;
}

// Extremes of |u| as integer keys.  |u| = sqrt((ux*ux + uy*uy) + uz*uz) is
// +0, positive, +inf or NaN (never -0: the squares are >= +0), so the key
// (bits << 1) + 1 (sign bit dropped) orders the values as the floats do,
// puts every NaN above +inf, and leaves 0 free for "no value"; one integer
// max or min per statistic replaces an FP64 max with its NaN handling.
constexpr uint32_t UKEY_INF = (0x7f800000u << 1) + 1u;
__device__ inline uint32_t umag_key(float um) { return (__float_as_uint(um) << 1) + 1u; }
__device__ inline double umag_of_key(uint32_t k) { return (double)__uint_as_float(k >> 1); }

// Wave reduction of a key statistic (max or min of keys) into sh[wid][S],
// decoded: an empty key gives the statistic's identity, a NaN key NaN.
template <int S, bool MIN>
__device__ inline void wave_key(uint32_t k, double (*sh)[POL_NSTAT + 1], int wid, int lane) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t t = __shfl_xor(k, o, 64);
    k = MIN ? min(k, t) : max(k, t);
  }
  if (lane == 0) {
    const bool empty = MIN ? k == ~0u : k == 0u;
    sh[wid][S] = empty ? stat_identity(S) : (k > UKEY_INF ? (double)NAN : umag_of_key(k));
  }
}

// Role A: rho and the velocity components.  Branch-free per cell: the
// non-solid statistics take selected values (the identity where a cell does
// not count), so mixed solid / fluid waves run one path.
template <typename R>
struct StatsA {
  double s_d = 0.0, s_d2 = 0.0, mx_d = -1.0e300, mom_x = 0.0, mom_y = 0.0, mom_z = 0.0, s_ux = 0.0, s_uy = 0.0,
         s_uz = 0.0, s_ux2 = 0.0, s_uy2 = 0.0, s_uz2 = 0.0, down = 0.0;
  R r_min = R(INFINITY), r_max = R(-INFINITY);  // over finite rho of non-solid cells (storage type: exact)
  // counts in 16-bit halves (a lane takes fewer than 65536 cells): cells |
  // non-solid << 16, NaN rho | infinite rho << 16
  uint32_t c_cells = 0, c_nonfinite = 0;
  __device__ void add(R r, R ux, R uy, R uz, uint32_t fl, bool outlet) {
    const double rd = (double)r;
    const double dux = (double)ux, duy = (double)uy, duz = (double)uz;
    const double d = rd - 1.0;
    s_d += d;
    s_d2 += d * d;
    mx_d = nanmax(mx_d, fabs(d));
    mom_x += rd * dux;
    mom_y += rd * duy;
    mom_z += rd * duz;
    s_ux += dux;
    s_uy += duy;
    s_uz += duz;
    s_ux2 += dux * dux;
    s_uy2 += duy * duy;
    s_uz2 += duz * duz;
    if (outlet) down += fmax(-duz, 0.0);  // block-uniform
    const bool ns = ((fl >> POL_KIND_SHIFT) & POL_KIND_MASK) != POL_KIND_SOLID;
    const bool nan = __builtin_isnan(r), inf = __builtin_isinf(r);
    c_cells += ns ? 0x10001u : 1u;
    c_nonfinite += (ns && nan) ? 1u : 0u;
    c_nonfinite += (ns && inf) ? 0x10000u : 0u;
    const bool fin = ns && __builtin_isfinite(r);
    r_min = fmin(r_min, fin ? r : R(INFINITY));
    r_max = fmax(r_max, fin ? r : R(-INFINITY));
  }
  __device__ void reduce(double (*sh)[POL_NSTAT + 1], int wid, int lane) {
    wave_stat<POL_STAT_CELLS>(0.0, c_cells & 0xffffu, sh, wid, lane);
    wave_stat<POL_STAT_SUM_DRHO>(s_d, 0u, sh, wid, lane);
    wave_stat<POL_STAT_SUM_DRHO2>(s_d2, 0u, sh, wid, lane);
    wave_stat<POL_STAT_MAX_ABS_DRHO>(mx_d, 0u, sh, wid, lane);
    wave_stat<POL_STAT_MOM_X>(mom_x, 0u, sh, wid, lane);
    wave_stat<POL_STAT_MOM_Y>(mom_y, 0u, sh, wid, lane);
    wave_stat<POL_STAT_MOM_Z>(mom_z, 0u, sh, wid, lane);
    wave_stat<POL_STAT_SUM_UX>(s_ux, 0u, sh, wid, lane);
    wave_stat<POL_STAT_SUM_UY>(s_uy, 0u, sh, wid, lane);
    wave_stat<POL_STAT_SUM_UZ>(s_uz, 0u, sh, wid, lane);
    wave_stat<POL_STAT_SUM_UX2>(s_ux2, 0u, sh, wid, lane);
    wave_stat<POL_STAT_SUM_UY2>(s_uy2, 0u, sh, wid, lane);
    wave_stat<POL_STAT_SUM_UZ2>(s_uz2, 0u, sh, wid, lane);
    wave_stat<POL_STAT_OUTLET_DOWN>(down, 0u, sh, wid, lane);
    wave_stat<POL_STAT_FLUID>(0.0, c_cells >> 16, sh, wid, lane);
    wave_stat<POL_STAT_FLUID_RHO_NAN>(0.0, c_nonfinite & 0xffffu, sh, wid, lane);
    wave_stat<POL_STAT_FLUID_RHO_INF>(0.0, c_nonfinite >> 16, sh, wid, lane);
    // +-inf only as the identity (the values are finite)
    const double lo = r_min == R(INFINITY) ? 1.0e300 : (double)r_min;
    const double hi = r_max == R(-INFINITY) ? -1.0e300 : (double)r_max;
    wave_stat<POL_STAT_FLUID_RHO_MIN>(lo, 0u, sh, wid, lane);
    wave_stat<POL_STAT_FLUID_RHO_MAX>(hi, 0u, sh, wid, lane);
    // the plane's density sum: sum (rho - 1) plus the cell count (exact),
    // taken from the two statistics just reduced
    if (lane == 0) sh[wid][POL_NSTAT] = sh[wid][POL_STAT_SUM_DRHO] + sh[wid][POL_STAT_CELLS];
  }
};

// Role B: |u| and what derives from it, and the temporal change.  Branch-free
// per cell like role A; the extremes of |u| are integer keys (umag_key).
// EXTRA: a previous state or a phase field is given (without one, the
// temporal change and the air mass are zero and every cell is water, and
// their accumulators are not kept).
template <bool EXTRA>
struct StatsB {
  double s_um = 0.0, s_um2 = 0.0, s_mach = 0.0, s_uma = 0.0, p_r2 = 0.0, p_u2 = 0.0, i_water = 0.0, i_air = 0.0,
         i_usum = 0.0;
  uint32_t k_max = 0u, k_fmax = 0u, k_imax = 0u, k_imin = ~0u;  // MAX_UMAG, FLUID_UMAX, INT_UMAX, INT_UMIN
  // counts in 16-bit halves (a lane takes fewer than 65536 cells):
  // Mach > 1e-10 | Mach > 0.1 << 16, |u| > 1e-10 | interior << 16,
  // NaN |u| | infinite |u| << 16 (non-solid cells)
  uint32_t c_mach = 0, c_um = 0, c_nonfinite = 0;
  // i: the cell's index within plane z (base = z nxy); in_box: 1 <= x, y, z
  // <= n - 2 (the interior test without the solid check)
  template <typename R>
  __device__ void add(const StatsArgs& a, size_t base, uint32_t i, bool in_box, R r, R ux, R uy, R uz, uint32_t fl) {
    const double rd = (double)r;
    const float umf = umag_f32(ux, uy, uz);
    const double um = (double)umf;
    const uint32_t key = umag_key(umf);
    s_um += um;
    s_um2 += um * um;
    k_max = max(k_max, key);
    // sums over Mach > 1e-10 and |u| > 1e-10 add +0 for the other cells
    // (a sum of non-negative values that starts at +0 is unchanged by it)
    const float mach = mach_of(umf);
    const bool act = mach > 1e-10f, ua = umf > 1e-10f;
    s_mach += act ? (double)mach : 0.0;
    c_mach += (act ? 1u : 0u) + (mach > 0.1f ? 0x10000u : 0u);
    s_uma += ua ? um : 0.0;
    bool water = true;
    if constexpr (EXTRA) {
      if (a.prev_rho) {
        const double e = rd - (double)((const R*)a.prev_rho)[base + i];
        p_r2 += e * e;
      }
      if (a.prev_u) {
        const R* pu = (const R*)a.prev_u + base + i;
        const double ex = (double)ux - (double)pu[0], ey = (double)uy - (double)pu[a.N],
                     ez = (double)uz - (double)pu[2 * (size_t)a.N];
        p_u2 += ex * ex + ey * ey + ez * ez;
      }
      if (a.phase) water = ((const R*)a.phase)[base + i] > R(0.1);
    }
    const bool ns = ((fl >> POL_KIND_SHIFT) & POL_KIND_MASK) != POL_KIND_SOLID;
    c_nonfinite += (ns && key > UKEY_INF) ? 1u : 0u;
    c_nonfinite += (ns && key == UKEY_INF) ? 0x10000u : 0u;
    k_fmax = max(k_fmax, (ns && key < UKEY_INF) ? key : 0u);
    const bool in = ns && in_box;
    c_um += (ua ? 1u : 0u) + (in ? 0x10000u : 0u);
    i_water += (in && water) ? rd : 0.0;
    if constexpr (EXTRA) i_air += (in && !water) ? rd : 0.0;
    i_usum += in ? um : 0.0;
    // interior max / min ignore NaN |u| (they keep +inf)
    const bool iv = in && key <= UKEY_INF;
    k_imax = max(k_imax, iv ? key : 0u);
    k_imin = min(k_imin, iv ? key : ~0u);
  }
  __device__ void reduce(double (*sh)[POL_NSTAT + 1], int wid, int lane) {
    wave_stat<POL_STAT_SUM_UMAG>(s_um, 0u, sh, wid, lane);
    wave_stat<POL_STAT_SUM_UMAG2>(s_um2, 0u, sh, wid, lane);
    wave_key<POL_STAT_MAX_UMAG, false>(k_max, sh, wid, lane);
    wave_stat<POL_STAT_SUM_MACH_ACTIVE>(s_mach, 0u, sh, wid, lane);
    wave_stat<POL_STAT_ACTIVE>(0.0, c_mach & 0xffffu, sh, wid, lane);
    wave_stat<POL_STAT_MACH_GT>(0.0, c_mach >> 16, sh, wid, lane);
    wave_stat<POL_STAT_DRHO_PREV2>(p_r2, 0u, sh, wid, lane);
    wave_stat<POL_STAT_DU_PREV2>(p_u2, 0u, sh, wid, lane);
    wave_stat<POL_STAT_FLUID_U_NAN>(0.0, c_nonfinite & 0xffffu, sh, wid, lane);
    wave_stat<POL_STAT_FLUID_U_INF>(0.0, c_nonfinite >> 16, sh, wid, lane);
    wave_key<POL_STAT_FLUID_UMAX, false>(k_fmax, sh, wid, lane);
    wave_stat<POL_STAT_INT_CELLS>(0.0, c_um >> 16, sh, wid, lane);
    wave_stat<POL_STAT_INT_WATER>(i_water, 0u, sh, wid, lane);
    wave_stat<POL_STAT_INT_AIR>(i_air, 0u, sh, wid, lane);
    wave_key<POL_STAT_INT_UMAX, false>(k_imax, sh, wid, lane);
    wave_key<POL_STAT_INT_UMIN, true>(k_imin, sh, wid, lane);
    wave_stat<POL_STAT_INT_USUM>(i_usum, 0u, sh, wid, lane);
    wave_stat<POL_STAT_SUM_UMAG_ACTIVE>(s_uma, 0u, sh, wid, lane);
    wave_stat<POL_STAT_UMAG_ACTIVE>(0.0, c_um & 0xffffu, sh, wid, lane);
  }
};

// rho, u and the flags of four consecutive cells (group j of a plane)
template <typename R>
struct Group4 {
  R r[4], x[4], y[4], w[4];
  uint32_t f[4];
  __device__ void load(const R* rho, const R* ux, const R* uy, const R* uz, const uint32_t* flp, uint32_t j,
                       uint32_t fluid_word) {
    load4(rho + 4 * (size_t)j, r);
    load4(ux + 4 * (size_t)j, x);
    load4(uy + 4 * (size_t)j, y);
    load4(uz + 4 * (size_t)j, w);
    if (flp) load4(flp + 4 * (size_t)j, f);
    else f[0] = f[1] = f[2] = f[3] = fluid_word;
  }
};

// Role loops over plane z: lane t of the pair takes cells (VEC: 4-cell
// groups) t, t + stride, ... of the plane.
template <typename R, bool VEC>
__device__ inline void stats_role_a(const StatsArgs& a, uint32_t z, uint32_t t0, uint32_t stride,
                                    double (*sh)[POL_NSTAT + 1], int wid, int lane) {
  StatsA<R> acc;
  const bool outlet = (int)z < a.z_out;
  const size_t base = (size_t)z * a.nxy;
  const R* rho = (const R*)a.rho + base;
  const R* ux_p = (const R*)a.u + base;
  const R* uy_p = ux_p + a.N;
  const R* uz_p = uy_p + a.N;
  const uint32_t* flp = a.flags ? a.flags + base : nullptr;
  const uint32_t fluid_word = (uint32_t)(POL_KIND_FLUID << POL_KIND_SHIFT);
  if constexpr (VEC) {
    for (uint32_t j = t0; j < a.nxy / 4; j += stride) {
      Group4<R> g;
      g.load(rho, ux_p, uy_p, uz_p, flp, j, fluid_word);
#pragma unroll
      for (int k = 0; k < 4; ++k) acc.add(g.r[k], g.x[k], g.y[k], g.w[k], g.f[k], outlet);
    }
  } else {
    for (uint32_t i = t0; i < a.nxy; i += stride)
      acc.add(rho[i], ux_p[i], uy_p[i], uz_p[i], flp ? flp[i] : fluid_word, outlet);
  }
  acc.reduce(sh, wid, lane);
}

template <typename R, bool VEC, bool EXTRA>
__device__ inline void stats_role_b(const StatsArgs& a, uint32_t z, uint32_t t0, uint32_t stride,
                                    double (*sh)[POL_NSTAT + 1], int wid, int lane) {
  StatsB<EXTRA> acc;
  const size_t base = (size_t)z * a.nxy;
  const R* rho = (const R*)a.rho + base;
  const R* ux_p = (const R*)a.u + base;
  const R* uy_p = ux_p + a.N;
  const R* uz_p = uy_p + a.N;
  const uint32_t* flp = a.flags ? a.flags + base : nullptr;
  const uint32_t fluid_word = (uint32_t)(POL_KIND_FLUID << POL_KIND_SHIFT);
  const bool zin = z >= 1 && z + 1 < a.nz;
  // 1 <= x <= nx - 2 and 1 <= y <= ny - 2 as unsigned range tests
  const uint32_t nx2 = a.nx - 2u, ny2 = a.ny - 2u;
  if constexpr (VEC) {
    // the launcher takes this form only for nx >= 4: a group's four cells
    // then span at most two rows
    for (uint32_t j = t0; j < a.nxy / 4; j += stride) {
      Group4<R> g;
      g.load(rho, ux_p, uy_p, uz_p, flp, j, fluid_word);
      const uint32_t y0 = fdiv(4 * j, a.dnx), x0 = 4 * j - y0 * a.nx;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const bool wrap = x0 + k >= a.nx;
        const uint32_t x = wrap ? x0 + k - a.nx : x0 + k, y = wrap ? y0 + 1 : y0;
        const bool in_box = zin && x - 1u < nx2 && y - 1u < ny2;
        acc.template add<R>(a, base, 4 * j + k, in_box, g.r[k], g.x[k], g.y[k], g.w[k], g.f[k]);
      }
    }
  } else {
    for (uint32_t i = t0; i < a.nxy; i += stride) {
      const uint32_t y = fdiv(i, a.dnx), x = i - y * a.nx;
      const bool in_box = zin && x - 1u < nx2 && y - 1u < ny2;
      acc.template add<R>(a, base, i, in_box, rho[i], ux_p[i], uy_p[i], uz_p[i], flp ? flp[i] : fluid_word);
    }
  }
  acc.reduce(sh, wid, lane);
}

// FP32 fields without a previous state or phase: at most 64 VGPRs (eight
// waves per SIMD to cover the load latency); the other forms hold more
// (FP64 fields twice the loaded bytes) and keep the compiler's allocation
template <typename R, bool VEC, bool EXTRA>
__global__ void __launch_bounds__(BLOCK, (sizeof(R) == 4 && !EXTRA) ? 8 : 1) k_field_stats(const StatsArgs a) {
  static_assert(BLOCK == 256, "two wave pairs per block");
  const uint32_t z = blockIdx.y;
  __shared__ double sh[BLOCK / 64][POL_NSTAT + 1];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint32_t t0 = blockIdx.x * (BLOCK / 2) + (threadIdx.x & (BLOCK / 2 - 1));
  const uint32_t stride = a.tiles * (BLOCK / 2);
  if (wid < 2) stats_role_a<R, VEC>(a, z, t0, stride, sh, wid, lane);
  else stats_role_b<R, VEC, EXTRA>(a, z, t0, stride, sh, wid, lane);
  __syncthreads();
  const size_t blk = (size_t)z * a.tiles + blockIdx.x;
  if (threadIdx.x < POL_NSTAT) {
    const int s = threadIdx.x;
    double r;
    switch (stat_role(s)) {
      case ROLE_A: r = stat_combine(s, sh[0][s], sh[1][s]); break;
      case ROLE_B: r = stat_combine(s, sh[2][s], sh[3][s]); break;
      case ROLE_REGION: {
        // the block's |u| sum, in the statistic of its plane's region
        const int region = (int)z >= a.z_top ? POL_STAT_SUM_UMAG_TOP
                           : (int)z >= a.z_mid ? POL_STAT_SUM_UMAG_MID
                                               : POL_STAT_SUM_UMAG_BOT;
        r = s == region ? sh[2][POL_STAT_SUM_UMAG] + sh[3][POL_STAT_SUM_UMAG] : 0.0;
        break;
      }
      default: r = stat_identity(s);
    }
    a.partial[blk * POL_NSTAT + s] = r;  // the block's statistics in one run of whole lines
  } else if (threadIdx.x == POL_NSTAT) {
    a.zpart[blk] = sh[0][POL_NSTAT] + sh[1][POL_NSTAT];
  }
}

// Fold the block partials: block s < POL_NSTAT folds statistic s, block
// POL_NSTAT + z the tiles of plane z.  Thread t combines the partials
// t, t + 256, ... in order, then the block combines its threads in a fixed
// tree: the result does not depend on timing (deterministic).
__global__ void __launch_bounds__(256) k_stats_final(const double* __restrict__ partial,
                                                     const double* __restrict__ zpart, uint32_t nblk, uint32_t tiles,
                                                     double* __restrict__ out, double* __restrict__ zout) {
  const uint32_t b = blockIdx.x;
  const bool stat = b < POL_NSTAT;
  const int s = stat ? (int)b : 0;
  double r = stat ? stat_identity(s) : 0.0;
  if (stat) {
    for (uint32_t k = threadIdx.x; k < nblk; k += 256) r = stat_combine(s, r, partial[(size_t)k * POL_NSTAT + s]);
  } else {
    const uint32_t z = b - POL_NSTAT;
    for (uint32_t k = threadIdx.x; k < tiles; k += 256) r += zpart[(size_t)z * tiles + k];
  }
  r = stat ? wave_combine(s, r) : wave_sum(r);
  __shared__ double sh[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = sh[0];
    for (int w = 1; w < 4; ++w) t = stat ? stat_combine(s, t, sh[w]) : t + sh[w];
    if (stat) out[s] = t;
    else zout[b - POL_NSTAT] = t;
  }
}

// Work item -> (dense cell, population slot, non-solid) for the bound launch
// list; the compact layout's slot comes from the chunk record.
template <bool SPARSE>
__device__ inline bool work_cell(const StepArgs& a, uint32_t lblk, uint32_t& cell, uint32_t& own, uint32_t& fl) {
  uint32_t yz, x, y, z;
  int rb[9];
  int sx = 0;
  bool in_run = false, plain = false, inside;  // plain: every flag word of the piece is 0
  if constexpr (SPARSE)
    inside = item_record(lblk * BLOCK + threadIdx.x, a.n_items, a.chunks, a.records, a.cshift, a.N, a.dnx, a.nx,
                         cell, y, z, x, rb, sx, in_run, plain);
  else
    inside = item_cell(lblk * BLOCK + threadIdx.x, a.n_items, a.chunks, a.cshift, a.N, a.dnx, a.nx, cell, yz, x);
  own = cell;
  if constexpr (SPARSE) own = slot_clamp(rb[4] + sx, a.P);
  fl = (inside && a.flags) ? (plain ? 0u : a.flags[cell]) : (inside ? (POL_KIND_FLUID << POL_KIND_SHIFT) : 0u);
  const uint32_t kind = (fl >> POL_KIND_SHIFT) & POL_KIND_MASK;
  return inside && kind != POL_KIND_SOLID;
}

// NaN / Inf populations of the current state over non-solid cells
// (numerical_stability.py:92-98); block partials {nan, inf}.
template <typename S, bool SPARSE>
__global__ void __launch_bounds__(BLOCK) k_pop_nonfinite(const StepArgs a, double* __restrict__ partial) {
  const uint32_t lblk = logical_block(a.per_xcd);
  if (lblk >= a.nblk) return;
  uint32_t cell, own, fl;
  double nn = 0.0, ni = 0.0;
  if (work_cell<SPARSE>(a, lblk, cell, own, fl)) {
    const S* __restrict__ f = (const S*)a.fin;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const double d = (double)Codec<S>::load(f + (size_t)q * a.P + own);
      nn += isnan(d) ? 1.0 : 0.0;
      ni += isinf(d) ? 1.0 : 0.0;
    }
  }
  __shared__ double sh[BLOCK / 64][2];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  nn = wave_sum(nn);
  ni = wave_sum(ni);
  if (lane == 0) {
    sh[wid][0] = nn;
    sh[wid][1] = ni;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double r = 0.0;
    for (int w = 0; w < BLOCK / 64; ++w) r += sh[w][threadIdx.x];
    partial[(size_t)lblk * 2 + threadIdx.x] = r;
  }
}

__global__ void __launch_bounds__(64) k_pop_final(const double* __restrict__ partial, uint32_t nblk,
                                                  double* __restrict__ out) {
  if (threadIdx.x < 2) {
    double r = 0.0;
    for (uint32_t b = 0; b < nblk; ++b) r += partial[(size_t)b * 2 + threadIdx.x];
    out[POL_STAT_FLUID_POP_NAN + threadIdx.x] = r;
  }
}

// Emergency stabilisation (numerical_stability.py:165-197) of non-solid cells.
template <typename S, bool SPARSE>
__global__ void __launch_bounds__(BLOCK) k_stabilize(const StepArgs a, float max_u, float min_rho, float max_rho) {
  using C = Codec<S>;
  using R = typename C::real;
  const uint32_t lblk = logical_block(a.per_xcd);
  if (lblk >= a.nblk) return;
  uint32_t cell, own, fl;
  if (!work_cell<SPARSE>(a, lblk, cell, own, fl)) return;
  R* __restrict__ rho = (R*)a.rho;
  R* __restrict__ u = (R*)a.u;
  const size_t N = a.N;
  R r = rho[cell];
  if (isnan(r) || isinf(r) || r <= R(0)) r = R(1);
  else if (r > (R)max_rho) r = (R)max_rho;
  else if (r < (R)min_rho) r = (R)min_rho;
  rho[cell] = r;
  R ux = u[cell], uy = u[N + cell], uz = u[2 * N + cell];
  const R um = (R)umag_f32(ux, uy, uz);
  if (isnan(um) || isinf(um)) {
    ux = uy = uz = R(0);
  } else if (um > (R)max_u) {
    const R s = (R)max_u / um;
    ux *= s;
    uy *= s;
    uz *= s;
  }
  u[cell] = ux;
  u[N + cell] = uy;
  u[2 * N + cell] = uz;
  // non-finite populations -> equilibrium of the repaired (rho, u)
  S* __restrict__ f = (S*)a.fout;
  const R usq = ux * ux + uy * uy + uz * uz;
  static_for<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    S* p = f + (size_t)q * a.P + own;
    const R d = C::load(p);
    if (isnan(d) || isinf(d)) C::store(p, deq_q<R, q>(CW<R>(q), r - R(1), r, ux, uy, uz, usq));
  });
}

// |u| and pressure fields of the export (enhanced_visualizer.py:1493-1505).
template <typename R>
__global__ void __launch_bounds__(256) k_export_fields(const R* __restrict__ rho, const R* __restrict__ u, uint32_t N,
                                                       R* __restrict__ mag, R* __restrict__ pres) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= N) return;
  const R ux = u[i], uy = u[(size_t)N + i], uz = u[2 * (size_t)N + i];
  if constexpr (sizeof(R) == 4) {
    mag[i] = umag_f32(ux, uy, uz);
    pres[i] = rho[i] * (float)(1.0 / 3.0);
  } else {
    mag[i] = sqrt((ux * ux + uy * uy) + uz * uz);
    pres[i] = rho[i] * (1.0 / 3.0);
  }
}

// ---- launchers ------------------------------------------------------------
template <typename R, bool VEC>
static void launch_stats(bool extra, dim3 grid, const StatsArgs& s, hipStream_t st) {
  if (extra) hipLaunchKernelGGL((k_field_stats<R, VEC, true>), grid, dim3(BLOCK), 0, st, s);
  else hipLaunchKernelGGL((k_field_stats<R, VEC, false>), grid, dim3(BLOCK), 0, st, s);
}

hipError_t launch_field_stats(bool real_double, const StatsArgs& s, double* out_dev, double* zout_dev,
                              hipStream_t st) {
  const dim3 grid(s.tiles, s.nz);
  // 16-byte loads need every plane (and so every u component) to start on a
  // 4-cell boundary; nx >= 4 keeps a 4-cell group within two rows
  const bool vec = s.nxy % 4 == 0 && s.nx >= 4;
  const bool extra = s.prev_rho || s.prev_u || s.phase;
  if (real_double) {
    if (vec) launch_stats<double, true>(extra, grid, s, st);
    else launch_stats<double, false>(extra, grid, s, st);
  } else {
    if (vec) launch_stats<float, true>(extra, grid, s, st);
    else launch_stats<float, false>(extra, grid, s, st);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_stats_final, dim3(POL_NSTAT + s.nz), dim3(256), 0, st, s.partial, s.zpart, s.tiles * s.nz,
                     s.tiles, out_dev, zout_dev);
  return hipGetLastError();
}

template <typename S>
static void launch_pop(const StepArgs& a, double* partial, dim3 grid, hipStream_t st) {
  if (a.records)
    hipLaunchKernelGGL((k_pop_nonfinite<S, true>), grid, dim3(BLOCK), 0, st, a, partial);
  else
    hipLaunchKernelGGL((k_pop_nonfinite<S, false>), grid, dim3(BLOCK), 0, st, a, partial);
}

hipError_t launch_pop_nonfinite(int storage, const StepArgs& a, double* partial, double* out_dev, const Grid& g,
                                hipStream_t st) {
  const dim3 grid(g.launch);
  switch (storage) {
    case POL_F32: launch_pop<float>(a, partial, grid, st); break;
    case POL_F16: launch_pop<_Float16>(a, partial, grid, st); break;
    case POL_BF16: launch_pop<bf16_t>(a, partial, grid, st); break;
    case POL_F64: launch_pop<double>(a, partial, grid, st); break;
    default: return hipErrorInvalidValue;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_pop_final, dim3(1), dim3(64), 0, st, partial, g.nblk, out_dev);
  return hipGetLastError();
}

template <typename S>
static void launch_stab(const StepArgs& a, float mu, float rmin, float rmax, dim3 grid, hipStream_t st) {
  if (a.records)
    hipLaunchKernelGGL((k_stabilize<S, true>), grid, dim3(BLOCK), 0, st, a, mu, rmin, rmax);
  else
    hipLaunchKernelGGL((k_stabilize<S, false>), grid, dim3(BLOCK), 0, st, a, mu, rmin, rmax);
}

hipError_t launch_stabilize(int storage, const StepArgs& a, float max_u, float min_rho, float max_rho, const Grid& g,
                            hipStream_t st) {
  const dim3 grid(g.launch);
  switch (storage) {
    case POL_F32: launch_stab<float>(a, max_u, min_rho, max_rho, grid, st); break;
    case POL_F16: launch_stab<_Float16>(a, max_u, min_rho, max_rho, grid, st); break;
    case POL_BF16: launch_stab<bf16_t>(a, max_u, min_rho, max_rho, grid, st); break;
    case POL_F64: launch_stab<double>(a, max_u, min_rho, max_rho, grid, st); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_export_fields(bool real_double, const void* rho, const void* u, uint32_t N, void* mag, void* pres,
                                hipStream_t st) {
  const dim3 grid((N + 255) / 256);
  if (real_double)
    hipLaunchKernelGGL(k_export_fields<double>, grid, dim3(256), 0, st, (const double*)rho, (const double*)u, N,
                       (double*)mag, (double*)pres);
  else
    hipLaunchKernelGGL(k_export_fields<float>, grid, dim3(256), 0, st, (const float*)rho, (const float*)u, N,
                       (float*)mag, (float*)pres);
  return hipGetLastError();
}

}  // namespace pol
