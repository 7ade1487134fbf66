// Device-side building blocks shared by the stream-collide, particle and
// adjoint kernels: D3Q19 tables, population storage codecs, index math and
// wave64 reductions.  gfx950 only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pourover_hip.h"

namespace pol {

// ---------------------------------------------------------------------------
// D3Q19 lattice in the order of config/core.py:36-47 (reference).  Opposite
// directions follow from e_opp = -e (the search of legacy/lbm_solver.py:431-439).
// ---------------------------------------------------------------------------
constexpr int Q = 19;
__host__ __device__ constexpr int EX(int q) {
  constexpr int t[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  return t[q];
}
__host__ __device__ constexpr int EY(int q) {
  constexpr int t[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
  return t[q];
}
__host__ __device__ constexpr int EZ(int q) {
  constexpr int t[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
  return t[q];
}
__host__ __device__ constexpr int OPP(int q) {
  constexpr int t[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
  return t[q];
}
// weights 1/3, 1/18, 1/36 and their exact integer multiples c_q = 36 w_q
template <typename R>
__host__ __device__ constexpr R W(int q) {
  return q == 0 ? R(1) / R(3) : (q < 7 ? R(1) / R(18) : R(1) / R(36));
}
template <typename R>
__host__ __device__ constexpr R CW(int q) {
  return q == 0 ? R(12) : (q < 7 ? R(2) : R(1));
}

// ---------------------------------------------------------------------------
// Population storage.  Every format stores the scaled deviation
// D_q = 36 (f_q - w_q);
// F16/BF16 are converted with round-to-nearest-even and widened to FP32 for
// all arithmetic.  load/store move deviations (no weight added here).
// ---------------------------------------------------------------------------
struct bf16_t {
  uint16_t bits;
};

template <typename S>
struct Codec;

template <>
struct Codec<float> {
  using real = float;
  __device__ static inline float decode(float v) { return v; }
  __device__ static inline float load(const float* p) { return *p; }
  __device__ static inline void store(float* p, float v) { *p = v; }
  __device__ static inline void store_nt(float* p, float v) { __builtin_nontemporal_store(v, p); }
};

template <>
struct Codec<double> {
  using real = double;
  __device__ static inline double decode(double v) { return v; }
  __device__ static inline double load(const double* p) { return *p; }
  __device__ static inline void store(double* p, double v) { *p = v; }
  __device__ static inline void store_nt(double* p, double v) { __builtin_nontemporal_store(v, p); }
};

// The FP16 codec also stores two values at once: one packed conversion
// (v_cvt_pk_f16_f32, round-to-nearest-even like the single one) and a
// low-half and a high-half 16-bit store.  BF16 keeps single conversions: the
// two-element form of its conversion (fptrunc of a float pair to a __bf16
// pair) is compiled wrongly by this toolchain (the second element is
// dropped and the first stored twice).
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));

template <>
struct Codec<_Float16> {
  using real = float;
  __device__ static inline float decode(_Float16 v) { return (float)v; }
  __device__ static inline float load(const _Float16* p) { return (float)(*p); }
  __device__ static inline void store(_Float16* p, float v) { *p = (_Float16)v; }
  __device__ static inline void store_nt(_Float16* p, float v) { __builtin_nontemporal_store((_Float16)v, p); }
  template <bool NT>
  __device__ static inline void store2(_Float16* p0, _Float16* p1, float a, float b) {
    const f16x2_t h = __builtin_convertvector((f32x2_t){a, b}, f16x2_t);
    if constexpr (NT) {
      __builtin_nontemporal_store(h.x, p0);
      __builtin_nontemporal_store(h.y, p1);
    } else {
      *p0 = h.x;
      *p1 = h.y;
    }
  }
};

__device__ inline float bf16_to_f32(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }
// round-to-nearest-even; NaN kept quiet
__device__ inline uint16_t f32_to_bf16(float v) {
  uint32_t u = __float_as_uint(v);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// gfx950 converts FP32 -> BF16 in one instruction (v_cvt_pk_bf16_f32,
// round-to-nearest-even, the rounding of f32_to_bf16 for every finite value)
__device__ inline uint16_t f32_to_bf16_hw(float v) { return __builtin_bit_cast(uint16_t, (__bf16)v); }

template <>
struct Codec<bf16_t> {
  using real = float;
  __device__ static inline float decode(bf16_t v) { return bf16_to_f32(v.bits); }
  __device__ static inline float load(const bf16_t* p) { return bf16_to_f32(p->bits); }
  __device__ static inline void store(bf16_t* p, float v) { p->bits = f32_to_bf16_hw(v); }
  __device__ static inline void store_nt(bf16_t* p, float v) { __builtin_nontemporal_store(f32_to_bf16_hw(v), &p->bits); }
};

// ---------------------------------------------------------------------------
// Exact unsigned division by a runtime constant (round-up multiplier,
// Granlund & Montgomery 1994, Thm 4.2): q = (n * m) >> (32 + l) with
// m = ceil(2^(32+l)/d), l = ceil(log2 d); exact for n < 2^32.  For d < 2^31
// and n < 2^31 the product n*m < 2^64 fits one 64-bit multiply.
// ---------------------------------------------------------------------------
struct FastDiv {
  uint64_t m;
  uint32_t shift;
  uint32_t d;
};
inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  f.shift = 32 + l;
  unsigned __int128 num = ((unsigned __int128)1) << (32 + l);
  f.m = (uint64_t)((num + d - 1) / d);
  return f;
}
__device__ __host__ inline uint32_t fdiv(uint32_t n, const FastDiv& f) {
  return (uint32_t)(((uint64_t)n * f.m) >> f.shift);
}

// ---------------------------------------------------------------------------
// Launch geometry: 1-D grid of 256-thread blocks over the x-fastest cell
// index (or the chunk list).  Hardware deals workgroups round-robin to the 8
// XCDs.  logical_block gives each XCD one contiguous run of blocks, so
// neighbouring rows stay in one L2 (the adjoint and diagnostics passes);
// ordered_block selects the order per launch (stream-collide), see there.
// ---------------------------------------------------------------------------
constexpr int BLOCK = 256;
constexpr int NXCD = 8;

struct Grid {
  uint32_t nblk;     // logical blocks covering the cells
  uint32_t per_xcd;  // ceil(nblk / 8)
  uint32_t launch;   // 8 * per_xcd physical blocks
};
inline Grid make_grid(uint64_t cells) {
  Grid g;
  g.nblk = (uint32_t)((cells + BLOCK - 1) / BLOCK);
  g.per_xcd = (g.nblk + NXCD - 1) / NXCD;
  g.launch = g.per_xcd * NXCD;
  return g;
}
__device__ inline uint32_t logical_block(uint32_t per_xcd) {
  const uint32_t b = blockIdx.x;
  return (b % NXCD) * per_xcd + b / NXCD;
}

// Block order of a launch.  The stream-collide step reads and writes 38
// population planes once each, so what counts is how many places of HBM the
// chip touches at once:
//   ORDER_LINEAR  block b works on logical block b: the whole chip sweeps one
//                 front through every plane (dense 224^3 periodic step
//                 0.306 ms, against 0.342-0.358 ms with XCD ranges, where 8
//                 fronts run at once);
//   k = 1..30     groups of 2^k consecutive logical blocks run on one XCD,
//                 inside super-blocks of 8 * 2^k blocks (one front; each XCD
//                 sees 2^k neighbouring blocks -- the compact layout's
//                 chunks of a row and their shared lines -- in its own L2);
//   ORDER_XCD_RANGES  logical_block above.
// Measured on the 224^3 V60 step (compact layout, streaming stores), FP32 /
// FP16 / FP64 ms: linear 0.1145 / 0.0623 / 0.2135, k = 2 0.1140 / 0.0609 /
// 0.2141, XCD ranges 0.1185 / 0.0655 / 0.2322; with 8-cell pieces and row
// records, FP32 / FP16: k = 2 0.1112 / 0.0620, linear 0.1119 / 0.0635,
// k = 1, 3, 4 0.1118-0.1122 / 0.0622-0.0628.  A step held in the
// last-level cache (128^3) is faster with XCD ranges: neighbour rows then
// hit the XCD's own L2.
constexpr uint32_t ORDER_LINEAR = 0;
constexpr uint32_t ORDER_XCD_RANGES = 31;
__device__ inline uint32_t ordered_block(uint32_t order, uint32_t per_xcd) {
  const uint32_t b = blockIdx.x;
  if (order == ORDER_LINEAR) return b;
  if (order == ORDER_XCD_RANGES) return (b % NXCD) * per_xcd + b / NXCD;
  const uint32_t G = 1u << order, S = NXCD * G;
  return (b / S) * S + (b % NXCD) * G + (b / NXCD) % G;
}
// Blocks to launch for an order: every logical block < g.nblk is covered;
// the ones past it exit at once.
inline uint32_t launch_blocks(uint32_t order, const Grid& g) {
  if (order == ORDER_LINEAR) return g.nblk;
  if (order == ORDER_XCD_RANGES) return g.launch;
  const uint32_t S = (uint32_t)NXCD << order;
  return (g.nblk + S - 1) / S * S;
}

// Plane-major buffer addressing.  Addr<true> keeps the element index
// q*N + i and its byte offset in 32 bits, so an access is one 64-bit SGPR base
// plus one 32-bit VGPR offset (global saddr form) instead of a 64-bit VGPR
// address pair; launchers select it when the whole buffer is below 4 GiB.
template <bool A32>
struct Addr;
template <>
struct Addr<true> {
  using idx = uint32_t;
  template <typename T>
  __device__ static T* at(T* b, uint32_t i) {
    return (T*)((char*)b + (uint32_t)(i * (uint32_t)sizeof(T)));
  }
  template <typename T>
  __device__ static const T* at(const T* b, uint32_t i) {
    return (const T*)((const char*)b + (uint32_t)(i * (uint32_t)sizeof(T)));
  }
};
template <>
struct Addr<false> {
  using idx = size_t;
  template <typename T>
  __device__ static T* at(T* b, size_t i) { return b + i; }
  template <typename T>
  __device__ static const T* at(const T* b, size_t i) { return b + i; }
};
// true when `planes` planes of n elements of `bytes` each stay below 4 GiB
inline bool fits_addr32(uint64_t planes, uint64_t n, uint64_t bytes) { return planes * n * bytes < (1ull << 32); }

// Work item -> cell.  Dense launches map item = cell.  Chunk launches map
// item to lane (item & (2^cshift - 1)) of row chunk chunks[item >> cshift]; a
// lane past the row end, or a cell index past the grid, is reported invalid,
// so no list content leads a kernel outside the buffers.
__device__ inline bool item_cell(uint32_t item, uint32_t n_items, const uint32_t* __restrict__ chunks, uint32_t cshift,
                                 uint32_t N, const FastDiv& dnx, uint32_t nx, uint32_t& cell, uint32_t& yz,
                                 uint32_t& x) {
  bool ok = item < n_items;
  const uint32_t lane = item & ((1u << cshift) - 1u);
  uint32_t c = 0;
  if (ok) c = chunks ? chunks[item >> cshift] + lane : item;
  ok = ok && c < N;
  if (!ok) c = 0;
  yz = fdiv(c, dnx);
  x = c - yz * nx;
  if (chunks && x < lane) ok = false;  // wrapped into the next row
  cell = ok ? c : 0u;
  if (!ok) {
    yz = 0;
    x = 0;
  }
  return ok;
}

// Work item -> cell for the compact layout: lane (item & (2^cshift - 1)) of
// the piece whose first dense cell is pieces[item >> cshift]; its row
// (z ny + y, one division by nx) selects the row record
// records[row * POL_RECORD_WORDS] (the row's first cell, the nine
// neighbour-row bases, the row's stored run, y | z << 16).  The cell index,
// x and the validity come from the piece alone, so a caller's load of the
// cell's flags goes out beside the row-record load.  Bit 31 of a piece word
// marks a plain piece (POL_PIECE_PLAIN: every cell of it inside the row has
// the flag word 0); plain returns it so that a caller may skip that load.
// rb[i] is the slot base of row (y + i%3 - 1, z + i/3 - 1).  sx = x0 + lane
// is the lane's x counted from its piece's row without wrapping (a stored run
// may extend past the row end as padding), and in_run tells whether sx lies
// in the own row's stored run, i.e. whether the lane owns slot rb[4] + sx.
__device__ inline bool item_record(uint32_t item, uint32_t n_items, const uint32_t* __restrict__ pieces,
                                   const uint32_t* __restrict__ records, uint32_t cshift, uint32_t N,
                                   const FastDiv& dnx, uint32_t nx, uint32_t& cell, uint32_t& y, uint32_t& z,
                                   uint32_t& x, int (&rb)[9], int& sx, bool& in_run, bool& plain) {
  const bool listed = item < n_items;
  const uint32_t lane = item & ((1u << cshift) - 1u);
  const uint32_t pw = listed ? pieces[item >> cshift] : 0u;
  const uint32_t first = pw & ~POL_PIECE_PLAIN;
  plain = (pw & POL_PIECE_PLAIN) != 0u;
  const uint32_t row = fdiv(first, dnx);  // z * ny + y
  const uint32_t x0 = first - row * nx;
  sx = (int)(x0 + lane);
  const uint32_t cc = first + lane;
  const bool ok = listed && x0 + lane < nx && cc < N;  // lanes past the row end are padding
  cell = ok ? cc : 0u;
  x = ok ? x0 + lane : 0u;
  uint32_t w[12] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
  if (listed && first < N) {  // row < ny nz: inside the record table
    const uint4* r = (const uint4*)(records + (size_t)row * POL_RECORD_WORDS);
    const uint4 a = r[0], b = r[1], c = r[2];
    w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
    w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
    w[8] = c.x; w[9] = c.y; w[10] = c.z; w[11] = c.w;
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) rb[i] = (int)w[1 + i];
  in_run = listed && first < N && sx >= (int)(w[10] & 0xffffu) && sx < (int)(w[10] >> 16);
  y = ok ? (w[11] & 0xffffu) : 0u;
  z = ok ? (w[11] >> 16) : 0u;
  return ok;
}

// Cell-flag bits 1..18: link q bounces back (bit 0, the rest population, never does).
constexpr uint32_t BOUNCE_BITS = 0x7fffeu;

// Slot of a compact-layout cell, clamped into the plane: no record content
// leads a kernel outside the population buffers.
__device__ inline uint32_t slot_clamp(int s, uint32_t P) { return min((uint32_t)s, P - 1u); }

// ---------------------------------------------------------------------------
// wave64 reductions (DPP/permute via shfl_xor) in FP64
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

// Per-block partials of the POL_NRED reduction slots.  Slots UMAX/RHOMAX
// reduce with max, RHOMIN with min, everything else with sum.
__device__ inline bool red_is_max(int s) { return s == POL_RED_UMAX || s == POL_RED_RHOMAX; }
__device__ inline bool red_is_min(int s) { return s == POL_RED_RHOMIN; }

template <int NW>
__device__ inline void block_reduce_store(double (&v)[POL_NRED], double* __restrict__ partial, uint32_t lblk) {
  __shared__ double sh[NW][POL_NRED];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int s = 0; s < POL_NRED; ++s) {
    double r = red_is_max(s) ? wave_max(v[s]) : (red_is_min(s) ? wave_min(v[s]) : wave_sum(v[s]));
    if (lane == 0) sh[wid][s] = r;
  }
  __syncthreads();
  if (threadIdx.x < POL_NRED) {
    const int s = threadIdx.x;
    double r = sh[0][s];
#pragma unroll
    for (int w = 1; w < NW; ++w) r = red_is_max(s) ? fmax(r, sh[w][s]) : (red_is_min(s) ? fmin(r, sh[w][s]) : r + sh[w][s]);
    partial[(size_t)lblk * POL_NRED + s] = r;
  }
}

__device__ inline double red_identity(int s) {
  return red_is_max(s) ? -1.0e300 : (red_is_min(s) ? 1.0e300 : 0.0);
}

}  // namespace pol
