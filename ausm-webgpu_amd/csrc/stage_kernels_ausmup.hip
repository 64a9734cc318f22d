// The fused RK-stage kernels of one scheme (SCHEME_AUSMUP): FP32 exact and fast
// math, compensated FP32 and FP64 (kernel templates in stage_kernels.hpp).
#include "stage_kernels.hpp"

namespace ausm {

hipError_t launch_stage_ausmup(const StageArgs& a, int fast, int ty, hipStream_t s) {
  return fast ? launch_ty<SCHEME_AUSMUP, true>(a, ty, s) : launch_ty<SCHEME_AUSMUP, false>(a, ty, s);
}

hipError_t launch_stage_comp_ausmup(const StageArgs& a, int fast, int ty, hipStream_t s) {
  return fast ? launch_ty_comp<SCHEME_AUSMUP, true>(a, ty, s) : launch_ty_comp<SCHEME_AUSMUP, false>(a, ty, s);
}

hipError_t launch_stage64_ausmup(const StageArgs64& a, int ty, hipStream_t s) {
  return launch_ty64<SCHEME_AUSMUP>(a, ty, s);
}

}  // namespace ausm
