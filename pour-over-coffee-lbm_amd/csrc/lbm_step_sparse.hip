// Compact-layout instantiations of the fused stream-collide kernel
// (stream_collide.h): populations stored per row run, chunk records give each
// thread its neighbour-row slot bases.  Compiled apart from the dense ones so
// the two sets build in parallel.
#include "stream_collide.h"

namespace pol {

// Streaming stores (a.stream_store, population traffic above the last-level
// cache) with the aligned runs: every chunk writes whole aligned segments,
// and in the one-front block order (ordered_block) neighbouring chunks
// write neighbouring segments at the same time, so no partial line is sent
// to HBM alone (224^3 V60, FP32 / FP16 / BF16 / FP64: 0.1140 / 0.0609 /
// 0.0680 / 0.2141 ms against 0.1225 / 0.0741 / 0.0735 / 0.2402 ms with
// cached stores and XCD ranges).  Runs packed without alignment meet at any
// offset; there streaming stores turn into read-modify-writes (0.216 ms
// against 0.143 ms), so such layouts should keep the step below the cache or
// run aligned.  XF selects the gather with the x-face wrap / clamp, for
// layouts with a run reaching x = 0 or nx - 1.
template <typename S, bool A32, bool XF>
static void launch_sparse(const StepArgs& a, bool store_f, dim3 grid, hipStream_t s) {
  dim3 block(BLOCK);
  if (store_f && a.stream_store) {
    hipLaunchKernelGGL((k_stream_collide<S, true, true, true, A32, true, true, XF>), grid, block, 0, s, a);
  } else if (store_f) {
    hipLaunchKernelGGL((k_stream_collide<S, true, true, true, A32, false, true, XF>), grid, block, 0, s, a);
  } else {
    hipLaunchKernelGGL((k_stream_collide<S, true, true, false, A32, false, true, XF>), grid, block, 0, s, a);
  }
}

template <typename S>
static hipError_t dispatch_sparse(const StepArgs& a, bool store_f, const Grid& g, hipStream_t s) {
  const bool a32 = fits_addr32(Q, a.P, sizeof(S));
  if (a32 && a.xface)
    launch_sparse<S, true, true>(a, store_f, dim3(launch_blocks(a.order, g)), s);
  else if (a32)
    launch_sparse<S, true, false>(a, store_f, dim3(launch_blocks(a.order, g)), s);
  else if (a.xface)
    launch_sparse<S, false, true>(a, store_f, dim3(launch_blocks(a.order, g)), s);
  else
    launch_sparse<S, false, false>(a, store_f, dim3(launch_blocks(a.order, g)), s);
  return hipGetLastError();
}

template <typename S, bool A32, bool XF>
static void launch_sparse_stream_only(const StepArgs& a, dim3 grid, hipStream_t s) {
  if (a.stream_store)
    hipLaunchKernelGGL((k_stream_collide<S, true, true, true, A32, true, true, XF, false>), grid, dim3(BLOCK), 0, s, a);
  else
    hipLaunchKernelGGL((k_stream_collide<S, true, true, true, A32, false, true, XF, false>), grid, dim3(BLOCK), 0, s, a);
}

template <typename S>
static hipError_t dispatch_sparse_stream_only(const StepArgs& a, const Grid& g, hipStream_t s) {
  const bool a32 = fits_addr32(Q, a.P, sizeof(S));
  const dim3 grid(launch_blocks(a.order, g));
  if (a32 && a.xface)
    launch_sparse_stream_only<S, true, true>(a, grid, s);
  else if (a32)
    launch_sparse_stream_only<S, true, false>(a, grid, s);
  else if (a.xface)
    launch_sparse_stream_only<S, false, true>(a, grid, s);
  else
    launch_sparse_stream_only<S, false, false>(a, grid, s);
  return hipGetLastError();
}

hipError_t launch_stream_only_sparse(int storage, const StepArgs& a, const Grid& g, hipStream_t s) {
  switch (storage) {
    case POL_F32: return dispatch_sparse_stream_only<float>(a, g, s);
    case POL_F16: return dispatch_sparse_stream_only<_Float16>(a, g, s);
    case POL_BF16: return dispatch_sparse_stream_only<bf16_t>(a, g, s);
    case POL_F64: return dispatch_sparse_stream_only<double>(a, g, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_stream_collide_sparse(int storage, const StepArgs& a, bool store_f, const Grid& g, hipStream_t s) {
  switch (storage) {
    case POL_F32: return dispatch_sparse<float>(a, store_f, g, s);
    case POL_F16: return dispatch_sparse<_Float16>(a, store_f, g, s);
    case POL_BF16: return dispatch_sparse<bf16_t>(a, store_f, g, s);
    case POL_F64: return dispatch_sparse<double>(a, store_f, g, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace pol
