// Internal kernel interface of libausm (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include "device_math.hpp"

namespace ausm {

constexpr int kWave = 64;
constexpr int kOwn = 60;           // columns a wavefront updates (2 halo lanes each side)
constexpr int kWavesPerBlock = 4;  // wavefronts per stage-kernel workgroup
// Partial wave-speed maxima a fresh stage 1 reduces itself: kDtAhead per lane
// are loaded at kernel entry, the rest (up to kWsInKernelMax) in a loop;
// beyond that the host reduces them once in a separate launch instead of in
// every wavefront (at 512 x 384 with one row per wavefront, 864 maxima, the
// in-kernel loop is the cheaper; at 1024 x 768 with two, 1632, the launch).
constexpr int kDtAhead = 8;
constexpr int kWsInKernelMax = 1024;

// One launch of the fused RK-stage kernel.  State buffers are SoA planes
// (rho, rho*u, rho*v, rho*E) of N interior rows with row pitch P and plane
// stride PS; vertices are two planes (x, y) of N+1 rows with pitch P.  R is
// the arithmetic type of the state, the vertices and the stage arithmetic
// (float; double for the all-FP64 solve).
template <typename R>
struct StageArgsT {
  const R* qin;   // stage input Q^n, Q1 or Q2 (read with a 2-cell halo)
  const R* qn;    // Q^n (read at the cell only; stages 2 and 3)
  R* qout;        // Q1, Q2 or Q^{n+1} (may be qn: written after its own read)
  const R* vx;
  const R* vy;
  const DevParamsT<R>* par;
  const R* ws_in;    // per-block max wave speeds of the last refresh
  R* ws_out;         // per-block max wave speeds of this stage's output (stage 3)
  double* res_out;   // sums of squares of the density residual per (row, strip), N x tiles_x (stage 1)
  R* maxws_cur;      // max wave speed the current step's dt is based on
  int n_ws_in;
  int M, N, P, PS;
  int stage;    // 0 = residual only (written to qout), 1, 2, 3
  int limiter;
  int fresh;    // stage 1: reduce ws_in into maxws_cur
  int refresh;  // stage 3: write ws_out
  int tiles_x;  // wavefront tiles around i
  int local_dt; // 1: each cell steps with its own CFL dt from Q^n (local time stepping)
  // compensated-state stages only: the increment accumulator (K1, then
  // K1 + K2) and the low part of the solution, both SoA planes like qout
  R* acc;
  R* lo;
  // limiter freezing (van Albada limiters, run-time limiter kernels only):
  // 0 off; 1 record psi and w of every face (stage 0 only); 2 reconstruct
  // with the record.  psi: 2 kPsiPerFace planes of psi_plane(N, P) values,
  // the 12 (psi_L, psi_R, w of rho, u, v, p) of the vertical faces (row jf =
  // face between rows jf-1 and jf, jf = 0..N), then the 12 of the
  // horizontal faces (row r = left face of the cell, r = 0..N-1)
  int freeze;
  R* psi;
};
using StageArgs = StageArgsT<float>;
using StageArgs64 = StageArgsT<double>;

// values per face of a limiter-freezing record, and the plane stride
constexpr int kPsiPerFace = 12;
__host__ __device__ inline size_t psi_plane(int N, int P) {
  return ((size_t)(N + 1) * (size_t)P + 63) / 64 * 64;
}

int stage_blocks(int M, int N, int ty);
int stage_tiles_x(int M);

// Host-side check of one stage launch: every pointer the stage dereferences
// is set and the sizes agree with the layout the kernel assumes.  The
// launchers refuse arguments that fail it (hipErrorInvalidValue) instead of
// letting the kernel read through a null pointer.  comp: compensated-state
// stages (which also use acc and lo).
template <typename R>
inline bool stage_args_ok(const StageArgsT<R>& a, bool comp) {
  if (!a.qin || !a.qout || !a.vx || !a.vy || !a.par) return false;
  if (a.stage < 0 || a.stage > 3) return false;
  if (a.M < 4 || a.N < 4 || a.P < a.M || a.P % kWave != 0) return false;
  if ((long long)a.PS < (long long)a.N * a.P || a.tiles_x != stage_tiles_x(a.M)) return false;
  if (a.stage >= 1) {
    if (!a.maxws_cur) return false;                                         // dt of the step
    if (a.stage == 1 && a.fresh && (!a.ws_in || a.n_ws_in < 1)) return false;  // fresh max
    if (a.stage >= 2 && !a.qn) return false;
    if (a.stage == 3 && a.refresh && !a.ws_out) return false;
    if (comp && (!a.acc || (a.stage == 3 && !a.lo))) return false;
  }
  if (a.freeze != 0) {
    if (a.freeze != 1 && a.freeze != 2) return false;
    if (!a.psi || (a.freeze == 1 && a.stage != 0)) return false;
    if (a.limiter != LIM_VANALBADA && a.limiter != LIM_VANALBADA_CLAMPED) return false;
  }
  return true;
}
hipError_t launch_stage(const StageArgs& a, int scheme, int fast, int ty, hipStream_t s);
// all-FP64 stage (exact arithmetic only); ty in {1, 2, 4}
hipError_t launch_stage64(const StageArgs64& a, int scheme, int ty, hipStream_t s);
// compensated-state FP32 stage; ty in {1, 2, 4}
hipError_t launch_stage_comp(const StageArgs& a, int scheme, int fast, int ty, hipStream_t s);

}  // namespace ausm
