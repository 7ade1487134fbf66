// Compact-layout instantiations of the fused stream-collide kernel
// (stream_collide.h): populations stored per row run, chunk records give each
// thread its neighbour-row slot bases.  Compiled apart from the dense ones so
// the two sets build in parallel.
#include "stream_collide.h"

namespace pol {

// Runs are packed without alignment, so the 128-byte lines at run boundaries
// are written in parts by neighbouring chunks; cached stores let L2 merge the
// parts (measured at 224^3 FP32: 288 MB written per step, the algorithmic
// 285 MB), while streaming stores send the parts to HBM as read-modify-writes
// (0.216 ms instead of 0.143 ms).  So no non-temporal variant here.
template <typename S, bool A32>
static void launch_sparse(const StepArgs& a, bool store_f, dim3 grid, hipStream_t s) {
  dim3 block(BLOCK);
  if (store_f) {
    hipLaunchKernelGGL((k_stream_collide<S, true, true, true, A32, false, true>), grid, block, 0, s, a);
  } else {
    hipLaunchKernelGGL((k_stream_collide<S, true, true, false, A32, false, true>), grid, block, 0, s, a);
  }
}

template <typename S>
static hipError_t dispatch_sparse(const StepArgs& a, bool store_f, const Grid& g, hipStream_t s) {
  if (fits_addr32(Q, a.P, sizeof(S)))
    launch_sparse<S, true>(a, store_f, dim3(g.launch), s);
  else
    launch_sparse<S, false>(a, store_f, dim3(g.launch), s);
  return hipGetLastError();
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
