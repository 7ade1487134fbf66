// Gathers of the compact row-run layout, shared by the forward stream-collide
// kernel and the adjoint kernel.
//
// Both walk the 19 links of one cell and hand each loaded value to a
// callback use(q, value) (q as std::integral_constant), so the forward kernel
// keeps the values and the adjoint reduces them as they arrive.
//
// Addressing.  A link's source (forward: x - e_q; reverse: x + e_p) lies in
// one of the nine rows (y +- 1, z +- 1) of the cell's chunk record, at column
// x -+ 1.  The row offset rb + x is bounded once per row (CompactRows): the
// layout's guard slots (slots 0, 1, P - 2 and P - 1 hold no cell) keep every
// legitimate source inside [2, P - 3] as seen from rb + x, so the bound only
// moves the reads of links whose value is replaced, and keeps them inside the
// plane.  The column shift -+ 1 is a constant address offset, so a link is
// one load on a scalar plane base with one 32-bit lane offset.  A link that reads the cell itself instead (forward: a bounced link
// q reads D*_{opp q}(x); reverse: the reader of f_p(x) is x's own bounce link
// opp p) switches the lane offset branch-free, with a flag bit as lane mask,
// to the own slot of the other plane: still one load, one memory round trip.
// Both links of a pair (q, opp q) address from the lower of the two planes,
// so every lane offset is non-negative.  XF (some run reaches x = 0 or
// nx - 1) adds the x-face wrap (periodic) or clamp of the dense gather.
#pragma once
#include <type_traits>

#include "cell_model.h"

namespace pol {

template <bool A32>
struct CompactRows {
  using AI = typename Addr<A32>::idx;  // byte offsets: 32-bit below 4 GiB of populations
  // ro[k]: byte offset of (row k's slot at this x) - 1, the slot rb[k] + x
  // clamped by one unsigned minimum: rb[k] + x - 1 >= 0 keeps its value
  // (every slot a kept value comes from is >= 2), anything else -- a
  // negative base of a row that has no run there, or past the plane -- lands
  // on P - 3.  Loads add the element back as part of their constant offset
  // (1 - e_x forward, 1 + e_x reverse), so a link stays one VGPR offset.
  AI ro[9];
  AI own_b;      // byte offset of the own slot, at least two elements
  int hi1;       // upper end of the x-face clamp range [2, P - 2]

  template <typename T>
  __device__ CompactRows(const int (&rb)[9], uint32_t x, uint32_t own, uint32_t PS, T*) {
    constexpr AI B = sizeof(T);
    hi1 = (int)PS - 2;
#pragma unroll
    for (int k = 0; k < 9; ++k) ro[k] = (AI)min((uint32_t)(rb[k] + (int)x - 1), PS - 3u) * B;
    own_b = (AI)(own < 2u ? 2u : own) * B;
  }
  __device__ int clampi(int s) const { return s < 2 ? 2 : (s > hi1 ? hi1 : s); }
};

// Forward (pull) sources of a cell: link q reads slot rb[(1 - e_y) + 3 (1 - e_z)]
// + x - e_x of plane q, or D*_{opp q} at its own slot when bit q of the cell
// flags is set.
// RAW: hand use() the stored value itself instead of its decoded real.
// SKIP: test per wave whether any lane has a bounce link (see below).
template <typename S, bool A32, bool XF, bool RAW = false, bool SKIP = (sizeof(S) == 2), typename Use>
__device__ __forceinline__ void compact_sources(const S* __restrict__ fin, uint32_t PS, const int (&rb)[9], uint32_t x,
                                                uint32_t nx, bool px, uint32_t own, uint32_t fl, Use&& use) {
  using C = Codec<S>;
  using AI = typename Addr<A32>::idx;
  using SI = std::conditional_t<A32, int32_t, int64_t>;
  constexpr AI B = sizeof(S);
  const AI PSB = (AI)PS * B;
  const CompactRows<A32> R(rb, x, own, PS, fin);
  bool xface = false;
  int oxm = -1, oxp = 1;
  if constexpr (XF) {
    xface = x == 0 || x == nx - 1;
    oxm = x == 0 ? (px ? (int)nx - 1 : 0) : -1;         // source at x-1
    oxp = x == nx - 1 ? (px ? -((int)nx - 1) : 0) : 1;  // source at x+1
  }
  auto links = [&](auto bounce_c) {
    static_for<0, Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int o = OPP(q);
      constexpr int lo = q < o ? q : o;
      constexpr int k = (1 - EY(q)) + 3 * (1 - EZ(q));
      AI off = R.ro[k];
      if constexpr (XF && EX(q) != 0) {
        if (xface) {  // wrapped / clamped source, folded back onto the constant 1 - e_x
          const int dx = EX(q) > 0 ? oxm : oxp;
          off = (AI)(R.clampi(rb[k] + (int)x + dx) + EX(q) - 1) * B;
        }
      }
      if constexpr (q != lo) off += (AI)(q - lo) * PSB;
      if constexpr (q != 0 && decltype(bounce_c)::value) {
        const AI bounce = R.own_b + (AI)(o - lo) * PSB + (AI)((SI)(EX(q) - 1) * (SI)B);
        const AI m = (AI)(SI)__builtin_amdgcn_sbfe((int)fl, q, 1);  // 0 or all ones
        off = (bounce & m) | (off & ~m);
      }
      const S* base = fin + (AI)lo * PS;
      const S v = *((const S*)((const char*)base + off) + (1 - EX(q)));
      if constexpr (RAW) use(qc, v);
      else use(qc, C::decode(v));
    });
  };
  // SKIP (default for 16-bit storage): waves whose cells have no bounce link
  // (about half of them in the V60 cup) skip the 18 address selections (a
  // wave-uniform test).  Measured on the 224^3 cup: BF16 step 0.0743 ->
  // 0.0713 ms, FP16 within 1 %; the 32- and 64-bit kernels, which run at
  // their access pattern's ceiling, the collisionless pass and the adjoint
  // lost 1-2 % with it, so they keep the single form.
  if constexpr (SKIP) {
    if (__ballot(fl & 0x7fffeu) != 0ull) links(std::true_type{});
    else links(std::false_type{});
  } else {
    links(std::true_type{});
  }
}

// Reverse (adjoint) readers of a cell's populations: f_p(x) is read either by
// x + e_p (normal link; bit p of the reverse flags rf) at slot
// rb[(1 + e_y) + 3 (1 + e_z)] + x + e_x of plane p, or by x itself through its
// bounce link opp p (slot own of plane opp p).  The value handed to use() is
// the one of the selected address; the caller zeroes it when neither reader
// exists.
template <typename T, bool A32, bool XF, typename Use>
__device__ __forceinline__ void compact_readers(const T* __restrict__ mu, uint32_t PS, const int (&rb)[9], uint32_t x,
                                                uint32_t nx, bool px, uint32_t own, uint32_t rf, Use&& use) {
  using AI = typename Addr<A32>::idx;
  using SI = std::conditional_t<A32, int32_t, int64_t>;
  constexpr AI B = sizeof(T);
  const AI PSB = (AI)PS * B;
  const CompactRows<A32> R(rb, x, own, PS, mu);
  bool xface = false;
  int oxm = -1, oxp = 1;
  if constexpr (XF) {
    xface = x == 0 || x == nx - 1;
    oxm = x == 0 ? (px ? (int)nx - 1 : 0) : -1;         // reader at x-1
    oxp = x == nx - 1 ? (px ? -((int)nx - 1) : 0) : 1;  // reader at x+1
  }
  static_for<0, Q>([&](auto pc) {
    constexpr int p = decltype(pc)::value;
    constexpr int o = OPP(p);
    constexpr int lo = p < o ? p : o;
    constexpr int k = (1 + EY(p)) + 3 * (1 + EZ(p));
    AI off = R.ro[k];
    if constexpr (XF && EX(p) != 0) {
      if (xface) {
        const int dx = EX(p) > 0 ? oxp : oxm;
        off = (AI)(R.clampi(rb[k] + (int)x + dx) - EX(p) - 1) * B;
      }
    }
    if constexpr (p != lo) off += (AI)(p - lo) * PSB;
    const AI bounce = R.own_b + (AI)(o - lo) * PSB - (AI)((SI)(EX(p) + 1) * (SI)B);
    const AI m = (AI)(SI)__builtin_amdgcn_sbfe((int)rf, p, 1);  // normal reader: 0 -> own slot
    off = (off & m) | (bounce & ~m);
    const T* base = mu + (AI)lo * PS;
    use(pc, *((const T*)((const char*)base + off) + (1 + EX(p))));
  });
}

}  // namespace pol
