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

// comp: compensated-state stages (every solve of the batch)
hipError_t launch_batch_stage(const BatchArgs& b, int fast, int comp, int ty, int blocks,
                              hipStream_t s);
hipError_t launch_batch_dt(const SolveDev* solves, const int* n_solves, int max_solves,
                           hipStream_t s);
hipError_t launch_batch_norm(const SolveDev* solves, const int* n_solves, int max_solves,
                             double* out, hipStream_t s);
hipError_t launch_batch_seed_ws(const float* part, int n_part, float* ws_part, int n_ws,
                                hipStream_t s);

}  // namespace ausm
