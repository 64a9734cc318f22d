// The persistent batched RK-stage kernel (templates): one launch per stage
// walks a (solve, tile) item list spanning every solve of a batch.  Compiled
// in one translation unit per arithmetic (batch_stage_<math>_<state>.hip).
#pragma once
#include "batch_kernels.hpp"
#include "reduce.hpp"
#include "stage_tile.hpp"

namespace ausm {

template <bool FAST, int TY, bool COMP, int LIM, int STAGE, bool FRZ = false>
__device__ __forceinline__ float batch_tile(int scheme, const DevParams& P, const StageArgs& a, int tx,
                                            int tyb, FaceLds<float>& s_face, float dt) {
  switch (scheme) {
    case SCHEME_SLAU2:
      return stage_tile<SCHEME_SLAU2, FAST, TY, float, COMP, LIM, STAGE, FRZ>(P, a, tx, tyb, s_face, dt);
    case SCHEME_SLAU:
      return stage_tile<SCHEME_SLAU, FAST, TY, float, COMP, LIM, STAGE, FRZ>(P, a, tx, tyb, s_face, dt);
    default:
      return stage_tile<SCHEME_AUSMUP, FAST, TY, float, COMP, LIM, STAGE, FRZ>(P, a, tx, tyb, s_face, dt);
  }
}

// STAGE: the launch's RK stage (1..3) at compile time, or -1 (the schedule
// tuner's stage-2-from-Q^n launches).  FRZ: the variant launched while any
// solve of the batch has its limiter frozen (the frozen solves' tiles take
// the freezing code; every other tile the code of the variant without it,
// whose register budget the freezing code would otherwise raise).
template <bool FAST, int TY, bool COMP, int STAGE, bool FRZ>
__global__ __launch_bounds__(256) void batch_stage_kernel(const BatchArgs b) {
  __shared__ FaceLds<float> s_face;  // one buffer for every inlined tile variant
  // workgroup g takes items [g n / B, (g+1) n / B): runs of floor or ceil
  // of n / B spread over every g, so every XCD's share is within one run of
  // n / 8 (fixed ceil(n / B) runs would leave the last workgroups -- the
  // last XCDs' -- without items whenever B does not divide n)
  const long long n = *b.n_items;
  const long long nb = gridDim.x;
  const long long g = xcd_remap(blockIdx.x, gridDim.x);
  const int lo = (int)(g * n / nb);
  const int hi = (int)((g + 1) * n / nb);
  for (int it = lo; it < hi; ++it) {
    const BatchItem item = b.items[it];
    const SolveDev* S = b.solves + item.solve;
    const StageArgs a = batch_args(*S, STAGE > 0 ? STAGE : b.stage, b.refresh, b.norm);
    const DevParams P = S->par;
    const int tx = item.tile % S->tiles_x;
    const int tyb = item.tile / S->tiles_x;
    const float dt = fminf(P.maxdt, P.cfl / *S->maxws_cur);
    // fast solves with van Albada (the reference's reconstruction, limiter
    // not frozen) or first order (no reconstruction) take code with the
    // limiter fixed at compile time
    float ws;
    if (FRZ && a.freeze)
      ws = batch_tile<FAST, TY, COMP, -1, STAGE, FRZ>(S->scheme, P, a, tx, tyb, s_face, dt);
    else if (FAST && a.limiter == LIM_VANALBADA)
      ws = batch_tile<FAST, TY, COMP, LIM_VANALBADA, STAGE>(S->scheme, P, a, tx, tyb, s_face, dt);
    else if (FAST && a.limiter == LIM_NONE)
      ws = batch_tile<FAST, TY, COMP, LIM_NONE, STAGE>(S->scheme, P, a, tx, tyb, s_face, dt);
    else
      ws = batch_tile<FAST, TY, COMP, -1, STAGE>(S->scheme, P, a, tx, tyb, s_face, dt);
    if (a.stage == 3 && a.refresh) block_ws_partial(ws, S->ws_part + item.tile);
  }
}

// per-solve max wave speed of the last refresh -> maxws_cur (one workgroup
// per table slot; inactive slots are skipped)
template <bool FAST, bool COMP, int STAGE, bool FRZ>
static hipError_t launch_bt_stage(const BatchArgs& b, int ty, int blocks, hipStream_t s) {
  switch (ty) {
    case 1: hipLaunchKernelGGL((batch_stage_kernel<FAST, 1, COMP, STAGE, FRZ>), dim3(blocks), dim3(256), 0, s, b); break;
    case 2: hipLaunchKernelGGL((batch_stage_kernel<FAST, 2, COMP, STAGE, FRZ>), dim3(blocks), dim3(256), 0, s, b); break;
    case 4: hipLaunchKernelGGL((batch_stage_kernel<FAST, 4, COMP, STAGE, FRZ>), dim3(blocks), dim3(256), 0, s, b); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <bool FAST, bool COMP, bool FRZ>
hipError_t launch_bt(const BatchArgs& b, int ty, int blocks, hipStream_t s) {
  switch (b.stage) {
    case 1: return launch_bt_stage<FAST, COMP, 1, FRZ>(b, ty, blocks, s);
    case 2: return launch_bt_stage<FAST, COMP, 2, FRZ>(b, ty, blocks, s);
    case 3: return launch_bt_stage<FAST, COMP, 3, FRZ>(b, ty, blocks, s);
    default: return launch_bt_stage<FAST, COMP, -1, FRZ>(b, ty, blocks, s);
  }
}

// one translation unit per (math, state): fast / exact x plain / compensated
hipError_t launch_bt_fast_plain(const BatchArgs& b, int frz, int ty, int blocks, hipStream_t s);
hipError_t launch_bt_fast_comp(const BatchArgs& b, int frz, int ty, int blocks, hipStream_t s);
hipError_t launch_bt_exact_plain(const BatchArgs& b, int frz, int ty, int blocks, hipStream_t s);
hipError_t launch_bt_exact_comp(const BatchArgs& b, int frz, int ty, int blocks, hipStream_t s);

}  // namespace ausm
