// Persistent batched stage kernels, exact math, compensated FP32 state.
#include "batch_stage.hpp"

namespace ausm {

hipError_t launch_bt_exact_comp(const BatchArgs& b, int frz, int ty, int blocks, hipStream_t s) {
  return frz ? launch_bt<false, true, true>(b, ty, blocks, s) : launch_bt<false, true, false>(b, ty, blocks, s);
}

}  // namespace ausm
