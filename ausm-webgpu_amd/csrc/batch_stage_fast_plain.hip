// Persistent batched stage kernels, fast math, plain FP32 state.
#include "batch_stage.hpp"

namespace ausm {

hipError_t launch_bt_fast_plain(const BatchArgs& b, int frz, int ty, int blocks, hipStream_t s) {
  return frz ? launch_bt<true, false, true>(b, ty, blocks, s) : launch_bt<true, false, false>(b, ty, blocks, s);
}

}  // namespace ausm
