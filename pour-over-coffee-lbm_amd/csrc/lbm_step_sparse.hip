// Compact-layout instantiations of the fused stream-collide kernel
// (stream_collide.h): populations stored per row run, chunk records give each
// thread its neighbour-row slot bases.  Compiled apart from the dense ones so
// the two sets build in parallel.
#include "stream_collide.h"

namespace pol {

// Neighbouring chunks share 128-byte lines (a 16-slot FP32 piece is half a
// line, packed runs meet at any offset); cached stores let L2 merge the parts
// (224^3 FP32: exactly the stored slots are written), while streaming stores
// send the parts to HBM as read-modify-writes (packed runs: 0.216 ms instead
// of 0.143 ms).  So no non-temporal variant here.  XF selects the gather with
// the x-face wrap / clamp, for layouts with a run reaching x = 0 or nx - 1.
template <typename S, bool A32, bool XF>
static void launch_sparse(const StepArgs& a, bool store_f, dim3 grid, hipStream_t s) {
  dim3 block(BLOCK);
  if (store_f) {
    hipLaunchKernelGGL((k_stream_collide<S, true, true, true, A32, false, true, XF>), grid, block, 0, s, a);
  } else {
    hipLaunchKernelGGL((k_stream_collide<S, true, true, false, A32, false, true, XF>), grid, block, 0, s, a);
  }
}

template <typename S>
static hipError_t dispatch_sparse(const StepArgs& a, bool store_f, const Grid& g, hipStream_t s) {
  const bool a32 = fits_addr32(Q, a.P, sizeof(S));
  if (a32 && a.xface)
    launch_sparse<S, true, true>(a, store_f, dim3(g.launch), s);
  else if (a32)
    launch_sparse<S, true, false>(a, store_f, dim3(g.launch), s);
  else if (a.xface)
    launch_sparse<S, false, true>(a, store_f, dim3(g.launch), s);
  else
    launch_sparse<S, false, false>(a, store_f, dim3(g.launch), s);
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
