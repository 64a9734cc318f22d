// Batched solves: one persistent fused RK-stage kernel walks a tile list
// that spans every grid in the batch.
//
//  * The solve table holds each solve's buffers, sizes, scheme, limiter and
//    free-stream parameters; the item list holds (solve, workgroup tile)
//    pairs sorted by scheme and solve, so a workgroup's run of items mostly
//    uses one scheme's code.
//  * The launch is persistent: a fixed number of workgroups per CU, each
//    taking one contiguous run of items (runs are dealt so that each XCD
//    gets neighbouring tiles and keeps their halos in its own L2).
//  * Per-solve dt comes from a small kernel that reduces each solve's
//    wave-speed maxima once per step (max is exact, so dt is the same
//    whatever else is in the batch); per-solve residual norms are reduced
//    from (row, strip) partials in a fixed order, so a solve's results are
//    bitwise the same alone and in any mix.
//  * Item count and solve flags live in device memory, so one captured
//    graph serves every batch composition.
#include "batch_kernels.hpp"
#include "reduce.hpp"
#include "stage_tile.hpp"

namespace ausm {

// one solve's tile with its scheme; LIM, STAGE and FRZ as in stage_tile
// (compile-time limiter / RK stage, or -1 for the run-time value; limiter
// freezing compiled in)
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
__global__ void batch_dt_kernel(const SolveDev* solves, const int* n_solves) {
  const int s = blockIdx.x;
  if (s >= *n_solves) return;
  const SolveDev& S = solves[s];
  if (!S.active) return;
  const float m = wave_max(strided_max(S.ws_part, S.n_ws, (int)threadIdx.x, 64));
  if (threadIdx.x == 0) *S.maxws_cur = m;
}

// per-solve FP64 residual norm (fixed-order reduction of the (row, strip)
// partials) into out[2*s], and into out[2*s+1] the larger of the max wave
// speed of the last step and that of the state it produced (NaN-propagating,
// so a solution that turned non-finite in the last step is caught now)
__global__ void batch_norm_kernel(const SolveDev* solves, const int* n_solves, double* out) {
  const int s = blockIdx.x;
  if (s >= *n_solves) return;
  const SolveDev& S = solves[s];
  if (!S.active) return;
  __shared__ double sm[256];
  __shared__ float smw[256];
  const int n = S.N * S.tiles_x;
  const double acc = strided_sum(S.res_part, n, (int)threadIdx.x, 256);
  const float w = strided_max(S.ws_part, S.n_ws, (int)threadIdx.x, 256);
  sm[threadIdx.x] = acc;
  smw[threadIdx.x] = w;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      sm[threadIdx.x] += sm[threadIdx.x + k];
      smw[threadIdx.x] = nan_max(smw[threadIdx.x], smw[threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * s] = sqrt(sm[0]);
    out[2 * s + 1] = (double)nan_max(*S.maxws_cur, smw[0]);
  }
}

// a solve's initial wave-speed partials in the stage-3 layout: the max of
// the per-block values in entry 0, zeros after it
__global__ void batch_seed_ws_kernel(const float* part, int n_part, float* ws_part, int n_ws) {
  const float m = wave_max(strided_max(part, n_part, (int)threadIdx.x, 64));
  for (int i = threadIdx.x; i < n_ws; i += 64) ws_part[i] = (i == 0) ? m : 0.0f;
}

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
static hipError_t launch_bt(const BatchArgs& b, int ty, int blocks, hipStream_t s) {
  switch (b.stage) {
    case 1: return launch_bt_stage<FAST, COMP, 1, FRZ>(b, ty, blocks, s);
    case 2: return launch_bt_stage<FAST, COMP, 2, FRZ>(b, ty, blocks, s);
    case 3: return launch_bt_stage<FAST, COMP, 3, FRZ>(b, ty, blocks, s);
    default: return launch_bt_stage<FAST, COMP, -1, FRZ>(b, ty, blocks, s);
  }
}

template <bool FRZ>
static hipError_t launch_bt_frz(const BatchArgs& b, int fast, int comp, int ty, int blocks,
                                hipStream_t s) {
  if (comp)
    return fast ? launch_bt<true, true, FRZ>(b, ty, blocks, s) : launch_bt<false, true, FRZ>(b, ty, blocks, s);
  return fast ? launch_bt<true, false, FRZ>(b, ty, blocks, s) : launch_bt<false, false, FRZ>(b, ty, blocks, s);
}

hipError_t launch_batch_stage(const BatchArgs& b, int fast, int comp, int frz, int ty, int blocks,
                              hipStream_t s) {
  return frz ? launch_bt_frz<true>(b, fast, comp, ty, blocks, s)
             : launch_bt_frz<false>(b, fast, comp, ty, blocks, s);
}

hipError_t launch_batch_dt(const SolveDev* solves, const int* n_solves, int max_solves,
                           hipStream_t s) {
  hipLaunchKernelGGL(batch_dt_kernel, dim3(max_solves), dim3(64), 0, s, solves, n_solves);
  return hipGetLastError();
}

hipError_t launch_batch_norm(const SolveDev* solves, const int* n_solves, int max_solves,
                             double* out, hipStream_t s) {
  hipLaunchKernelGGL(batch_norm_kernel, dim3(max_solves), dim3(256), 0, s, solves, n_solves, out);
  return hipGetLastError();
}

hipError_t launch_batch_seed_ws(const float* part, int n_part, float* ws_part, int n_ws,
                                hipStream_t s) {
  hipLaunchKernelGGL(batch_seed_ws_kernel, dim3(1), dim3(64), 0, s, part, n_part, ws_part, n_ws);
  return hipGetLastError();
}

}  // namespace ausm
