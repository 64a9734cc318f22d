// Batched-solve device structures and launchers (internal).
#pragma once
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "kernels.hpp"

namespace ausm {

// one solve of the batch, as the kernels see it
struct SolveDev {
  DevParams par;
  float* q[3];  // Q^n, Q1, Q2 (SoA planes)
  const float* vx;
  const float* vy;
  float* ws_part;    // per-workgroup-tile max wave speeds (stage-3 layout)
  double* res_part;  // density-residual sums of squares per (row, strip)
  float* maxws_cur;  // max wave speed of this step's dt
  float* acc;        // compensated batches: increment accumulator planes
  float* lo;         // compensated batches: low part of the solution
  int M, N, P, PS, tiles_x;
  int scheme, limiter;
  int n_ws;    // workgroup tiles of this solve = entries of ws_part
  int active;  // 0: slot not stepped
  int local_dt;  // 1: local time stepping
  int freeze;    // 2: limiter frozen (psi recorded), else 0
  float* psi;    // psi planes (solves that may freeze)
};

struct BatchItem {
  int solve;  // index into the solve table
  int tile;   // workgroup tile: tx + tiles_x * tyb
};

struct BatchArgs {
  const SolveDev* solves;
  const BatchItem* items;
  const int* n_items;  // device-resident so that captured graphs serve any mix
  int stage;           // 1, 2, 3; 4 = schedule tuning (no state change)
  int refresh;
  int norm;            // stage 1: write the residual-norm partials
};

// The stage arguments of one solve of the batch: the persistent kernel
// builds them per item, and the host's initial-residual pass of a newly
// admitted solve builds its stage-1 launch from the same function.  par
// is left null: the batch kernel passes each solve's parameters by value.
__host__ __device__ inline StageArgs batch_args(const SolveDev& S, int stage, int refresh,
                                                int norm) {
  StageArgs a{};
  a.vx = S.vx;
  a.vy = S.vy;
  a.par = nullptr;
  a.ws_in = S.ws_part;
  a.ws_out = S.ws_part;
  a.res_out = norm ? S.res_part : nullptr;
  a.maxws_cur = S.maxws_cur;
  a.n_ws_in = S.n_ws;
  a.M = S.M;
  a.N = S.N;
  a.P = S.P;
  a.PS = S.PS;
  a.limiter = S.limiter;
  a.tiles_x = S.tiles_x;
  a.local_dt = S.local_dt;
  a.acc = S.acc;
  a.lo = S.lo;
  a.freeze = S.freeze;
  a.psi = S.psi;
  a.fresh = 0;
  a.refresh = refresh;
  if (stage == 1) {
    a.stage = 1;
    a.qin = S.q[0];
    a.qn = S.q[0];
    a.qout = S.q[1];
  } else if (stage == 2) {
    a.stage = 2;
    a.qin = S.q[1];
    a.qn = S.q[0];
    a.qout = S.q[2];
  } else if (stage == 3) {
    a.stage = 3;
    a.qin = S.q[2];
    a.qn = S.q[0];
    a.qout = S.q[0];
  } else {
    // schedule tuning: stage-2 arithmetic from Q^n into the scratch Q2
    a.stage = 2;
    a.qin = S.q[0];
    a.qn = S.q[0];
    a.qout = S.q[2];
    a.refresh = 0;
  }
  return a;
}

// comp: compensated-state stages (every solve of the batch); frz: some
// solve of the batch has its limiter frozen
hipError_t launch_batch_stage(const BatchArgs& b, int fast, int comp, int frz, int ty, int blocks,
                              hipStream_t s);
hipError_t launch_batch_dt(const SolveDev* solves, const int* n_solves, int max_solves,
                           hipStream_t s);
hipError_t launch_batch_norm(const SolveDev* solves, const int* n_solves, int max_solves,
                             double* out, hipStream_t s);
hipError_t launch_batch_seed_ws(const float* part, int n_part, float* ws_part, int n_ws,
                                hipStream_t s);

}  // namespace ausm
