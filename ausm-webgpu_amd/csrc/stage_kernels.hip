// Stage-kernel launch entry points: argument checks and the dispatch over
// schemes.  The kernels (stage_kernels.hpp) are compiled in one translation
// unit per scheme (stage_kernels_slau2.hip, _slau.hip, _ausmup.hip), so that
// the build compiles them in parallel.
#include "stage_kernels.hpp"

namespace ausm {

int stage_tiles_x(int M) { return (M + kOwn - 1) / kOwn; }

int stage_blocks(int M, int N, int ty) {
  const int rows_per_block = ty * kWavesPerBlock;
  return stage_tiles_x(M) * ((N + rows_per_block - 1) / rows_per_block);
}

hipError_t launch_stage(const StageArgs& a, int scheme, int fast, int ty, hipStream_t s) {
  if (!stage_args_ok(a, false)) return hipErrorInvalidValue;
  switch (scheme) {
    case SCHEME_SLAU2: return launch_stage_slau2(a, fast, ty, s);
    case SCHEME_SLAU: return launch_stage_slau(a, fast, ty, s);
    case SCHEME_AUSMUP: return launch_stage_ausmup(a, fast, ty, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_stage_comp(const StageArgs& a, int scheme, int fast, int ty, hipStream_t s) {
  if (!stage_args_ok(a, true)) return hipErrorInvalidValue;
  switch (scheme) {
    case SCHEME_SLAU2: return launch_stage_comp_slau2(a, fast, ty, s);
    case SCHEME_SLAU: return launch_stage_comp_slau(a, fast, ty, s);
    case SCHEME_AUSMUP: return launch_stage_comp_ausmup(a, fast, ty, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_stage64(const StageArgs64& a, int scheme, int ty, hipStream_t s) {
  if (!stage_args_ok(a, false)) return hipErrorInvalidValue;
  switch (scheme) {
    case SCHEME_SLAU2: return launch_stage64_slau2(a, ty, s);
    case SCHEME_SLAU: return launch_stage64_slau(a, ty, s);
    case SCHEME_AUSMUP: return launch_stage64_ausmup(a, ty, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace ausm
