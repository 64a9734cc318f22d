// Fused TVD-RK3 stage kernel.
//
// One launch replaces the reference's five passes of a stage (boundary,
// vertical flux, horizontal flux, residual, stage update:
// src/main.js:744-776, src/shaders/ausm.js:23-528):
//   * a workgroup is four wavefronts stacked in j; each wavefront owns a
//     strip of 60 cells around i (64 lanes, two halo lanes each side) and TY
//     rows up j, and loads its whole window (rows j0-2 .. j0+TY+1; above the
//     workgroup's bottom wavefront from j0-1, since row j0-2 only feeds the
//     bottom face, which those wavefronts take from LDS) before any
//     arithmetic;
//   * each window row is converted to primitives once, and the differences
//     of neighbouring rows that MUSCL works on are formed once and shared by
//     the faces that use them;
//   * i-neighbours (MUSCL stencil, the right face's flux, the i+1 vertex)
//     come from the adjacent lanes by DPP wavefront shifts;
//   * a wavefront's top face is the bottom face of the wavefront above: it
//     is computed once and passed up through LDS;
//   * ghost rows are never stored: the body and far-field ghosts are formed
//     on the fly from interior rows 1 and N when the window reaches them;
//   * each face flux is evaluated once, with its length; the residual and
//     the RK update follow in registers;
//   * stage 3 also forms the CFL wave speed of the new state (ausm.js:617-630)
//     and a per-workgroup max (no atomics); the next stage 1 reduces those
//     maxima itself (loaded at entry, reduced after the first face), so dt
//     needs no extra launch and is bitwise deterministic;
//   * stage 1 writes FP64 sums of the squared density residual per (row,
//     60-column strip) for the convergence norm, so the norm does not depend
//     on the tile shape.
// Memory per cell-step: state 8 x 16 B streams + one vertex per stage.
#pragma once
#include "stage_tile.hpp"

namespace ausm {

// LIM: the limiter as a compile-time constant (the fast path's van Albada,
// the reference's own reconstruction), or -1 for the run-time choice.
// STAGE: the RK stage as a compile-time constant (1..3), or -1 for a.stage.
// WPE: the wavefronts per SIMD the register allocation must allow (the
// exact-math van Albada stages of one and two rows fit 4 in 128 VGPRs
// without spilling, against 3 at their unconstrained 140-170; every other
// variant is left to the compiler).  FRZ: limiter freezing compiled in
// (launched for frozen solves and recording passes only).
template <int SCHEME, bool FAST, int TY, int LIM = -1, int STAGE = -1, int WPE = 1, bool FRZ = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void stage_kernel(
    const StageArgs a) {
  const DevParams P = *a.par;
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = tile % a.tiles_x;
  const int tyb = tile / a.tiles_x;
  const int stage = STAGE >= 0 ? STAGE : a.stage;
  const DtAhead<float> ahead = stage_dt_load(a, stage);
  __shared__ FaceLds<float> s_face;
  const float ws_acc = stage_tile<SCHEME, FAST, TY, float, false, LIM, STAGE, FRZ>(
      P, a, tx, tyb, s_face, [&] { return stage_dt_finish(P, a, blockIdx.x == 0, ahead, stage); });
  if (stage == 3 && a.refresh) block_ws_partial(ws_acc, a.ws_out + blockIdx.x);
}

// All-FP64 variant (exact arithmetic): the same tile walk over double
// state and vertex planes, for the FP64 comparison solve.
template <int SCHEME, int TY, int LIM = -1, int STAGE = -1, bool FRZ = false>
__global__ __launch_bounds__(256) void stage_kernel64(const StageArgs64 a) {
  const DevParamsT<double> P = *a.par;
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = tile % a.tiles_x;
  const int tyb = tile / a.tiles_x;
  const int stage = STAGE >= 0 ? STAGE : a.stage;
  const DtAhead<double> ahead = stage_dt_load(a, stage);
  __shared__ FaceLds<double> s_face;
  const double ws_acc = stage_tile<SCHEME, false, TY, double, false, LIM, STAGE, FRZ>(
      P, a, tx, tyb, s_face, [&] { return stage_dt_finish(P, a, blockIdx.x == 0, ahead, stage); });
  if (stage == 3 && a.refresh) block_ws_partial(ws_acc, a.ws_out + blockIdx.x);
}

// Compensated-state FP32 variant (the COMP stages of stage_tile).
template <int SCHEME, bool FAST, int TY, int LIM = -1, int STAGE = -1, bool FRZ = false>
__global__ __launch_bounds__(256) void stage_kernel_comp(const StageArgs a) {
  const DevParams P = *a.par;
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = tile % a.tiles_x;
  const int tyb = tile / a.tiles_x;
  const int stage = STAGE >= 0 ? STAGE : a.stage;
  const DtAhead<float> ahead = stage_dt_load(a, stage);
  __shared__ FaceLds<float> s_face;
  const float ws_acc = stage_tile<SCHEME, FAST, TY, float, true, LIM, STAGE, FRZ>(
      P, a, tx, tyb, s_face, [&] { return stage_dt_finish(P, a, blockIdx.x == 0, ahead, stage); });
  if (stage == 3 && a.refresh) block_ws_partial(ws_acc, a.ws_out + blockIdx.x);
}

template <int SCHEME, bool FAST, int LIM, int STAGE, bool FRZ = false>
static hipError_t launch_ty_lim(const StageArgs& a, int ty, hipStream_t s) {
  const int nb = stage_blocks(a.M, a.N, ty);
  constexpr int W = (!FAST && LIM == LIM_VANALBADA && STAGE > 0) ? 4 : 1;
  switch (ty) {
    case 1: hipLaunchKernelGGL((stage_kernel<SCHEME, FAST, 1, LIM, STAGE, W, FRZ>), dim3(nb), dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((stage_kernel<SCHEME, FAST, 2, LIM, STAGE, W, FRZ>), dim3(nb), dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL((stage_kernel<SCHEME, FAST, 4, LIM, STAGE, 1, FRZ>), dim3(nb), dim3(256), 0, s, a); break;
    case 8: hipLaunchKernelGGL((stage_kernel<SCHEME, FAST, 8, LIM, STAGE, 1, FRZ>), dim3(nb), dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// van Albada stages 1-3 (the stepping path of the reference's
// reconstruction, fast and exact math) run code specialised on limiter and
// stage; everything else (other limiters, residual evaluation, limiter
// freezing) the general kernel
template <int SCHEME, bool FAST>
static hipError_t launch_ty(const StageArgs& a, int ty, hipStream_t s) {
  if (a.freeze) return launch_ty_lim<SCHEME, FAST, -1, -1, true>(a, ty, s);
  if (a.limiter == LIM_VANALBADA) {
    switch (a.stage) {
      case 1: return launch_ty_lim<SCHEME, FAST, LIM_VANALBADA, 1>(a, ty, s);
      case 2: return launch_ty_lim<SCHEME, FAST, LIM_VANALBADA, 2>(a, ty, s);
      case 3: return launch_ty_lim<SCHEME, FAST, LIM_VANALBADA, 3>(a, ty, s);
      default: break;
    }
  }
  return launch_ty_lim<SCHEME, FAST, -1, -1>(a, ty, s);
}

template <int SCHEME, bool FAST, int LIM, int STAGE, bool FRZ = false>
static hipError_t launch_ty_comp_lim(const StageArgs& a, int ty, hipStream_t s) {
  const int nb = stage_blocks(a.M, a.N, ty);
  switch (ty) {
    case 1: hipLaunchKernelGGL((stage_kernel_comp<SCHEME, FAST, 1, LIM, STAGE, FRZ>), dim3(nb), dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((stage_kernel_comp<SCHEME, FAST, 2, LIM, STAGE, FRZ>), dim3(nb), dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL((stage_kernel_comp<SCHEME, FAST, 4, LIM, STAGE, FRZ>), dim3(nb), dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// fast van Albada stages 1-3 specialised as in launch_ty
template <int SCHEME, bool FAST>
static hipError_t launch_ty_comp(const StageArgs& a, int ty, hipStream_t s) {
  if (a.freeze) return launch_ty_comp_lim<SCHEME, FAST, -1, -1, true>(a, ty, s);
  if constexpr (FAST) {
    if (a.limiter == LIM_VANALBADA) {
      switch (a.stage) {
        case 1: return launch_ty_comp_lim<SCHEME, FAST, LIM_VANALBADA, 1>(a, ty, s);
        case 2: return launch_ty_comp_lim<SCHEME, FAST, LIM_VANALBADA, 2>(a, ty, s);
        case 3: return launch_ty_comp_lim<SCHEME, FAST, LIM_VANALBADA, 3>(a, ty, s);
        default: break;
      }
    }
  }
  return launch_ty_comp_lim<SCHEME, FAST, -1, -1>(a, ty, s);
}

template <int SCHEME, int LIM, int STAGE, bool FRZ = false>
static hipError_t launch_ty64_lim(const StageArgs64& a, int ty, hipStream_t s) {
  const int nb = stage_blocks(a.M, a.N, ty);
  switch (ty) {
    case 1: hipLaunchKernelGGL((stage_kernel64<SCHEME, 1, LIM, STAGE, FRZ>), dim3(nb), dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((stage_kernel64<SCHEME, 2, LIM, STAGE, FRZ>), dim3(nb), dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL((stage_kernel64<SCHEME, 4, LIM, STAGE, FRZ>), dim3(nb), dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// van Albada stages 1-3 specialised as in launch_ty
template <int SCHEME>
static hipError_t launch_ty64(const StageArgs64& a, int ty, hipStream_t s) {
  if (a.freeze) return launch_ty64_lim<SCHEME, -1, -1, true>(a, ty, s);
  if (a.limiter == LIM_VANALBADA) {
    switch (a.stage) {
      case 1: return launch_ty64_lim<SCHEME, LIM_VANALBADA, 1>(a, ty, s);
      case 2: return launch_ty64_lim<SCHEME, LIM_VANALBADA, 2>(a, ty, s);
      case 3: return launch_ty64_lim<SCHEME, LIM_VANALBADA, 3>(a, ty, s);
      default: break;
    }
  }
  return launch_ty64_lim<SCHEME, -1, -1>(a, ty, s);
}

// per-scheme launchers, one translation unit each (stage_kernels_<scheme>.hip)
#define AUSM_SCHEME_LAUNCHERS(name)                                                          \
  hipError_t launch_stage_##name(const StageArgs& a, int fast, int ty, hipStream_t s);      \
  hipError_t launch_stage_comp_##name(const StageArgs& a, int fast, int ty, hipStream_t s); \
  hipError_t launch_stage64_##name(const StageArgs64& a, int ty, hipStream_t s);
AUSM_SCHEME_LAUNCHERS(slau2)
AUSM_SCHEME_LAUNCHERS(slau)
AUSM_SCHEME_LAUNCHERS(ausmup)
#undef AUSM_SCHEME_LAUNCHERS

}  // namespace ausm
