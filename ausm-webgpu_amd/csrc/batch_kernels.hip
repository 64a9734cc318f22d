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
#include "batch_stage.hpp"

namespace ausm {

// one solve's tile with its scheme; LIM, STAGE and FRZ as in stage_tile
// (compile-time limiter / RK stage, or -1 for the run-time value; limiter
// freezing compiled in)
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

hipError_t launch_batch_stage(const BatchArgs& b, int fast, int comp, int frz, int ty, int blocks,
                              hipStream_t s) {
  if (fast) return comp ? launch_bt_fast_comp(b, frz, ty, blocks, s) : launch_bt_fast_plain(b, frz, ty, blocks, s);
  return comp ? launch_bt_exact_comp(b, frz, ty, blocks, s) : launch_bt_exact_plain(b, frz, ty, blocks, s);
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
