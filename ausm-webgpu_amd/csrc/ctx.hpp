// Single-instance solver context (internal).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <vector>

#include "device_math.hpp"
#include "kernels.hpp"

namespace ausm {

struct GraphEntry {
  int ty, steps, n_ws_in;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

}  // namespace ausm

struct ausm_ctx {
  int device = 0;
  int M = 0, N = 0, P = 0, PS = 0;
  hipStream_t stream = nullptr;
  ausm::DevParams hpar{};
  ausm::DevParams* dpar = nullptr;
  float* q[3] = {nullptr, nullptr, nullptr};  // Q^n, Q1, Q2 (SoA planes)
  float* vx = nullptr;                        // FP32 vertices used by the solver
  float* vy = nullptr;
  double* g64[4] = {nullptr, nullptr, nullptr, nullptr};  // FP64 grid ping-pong (x,y,x',y')
  int g_cur = -1;                                          // 0 or 2: which pair holds the grid
  float* ws_part = nullptr;   // per-workgroup max wave speeds
  double* res_part = nullptr; // density-residual sums of squares per (row, 60-column strip)
  float* scal_f = nullptr;    // [0] max wave speed of the current dt, [1..] scratch
  double* scal_d = nullptr;   // [0] residual norm, [2..5] forces
  float* stage_buf = nullptr; // AoS staging, (N+3) x M x 4 floats
  double* seg = nullptr;      // per-segment forces (2M) + cp (M)
  void* pinned = nullptr;     // small pinned read-back area
  int max_parts = 0;
  int n_ws_valid = 0;
  bool ws_fresh = false;
  bool has_grid = false, has_state = false;
  int scheme = 0, limiter = 0, math = 0, refresh = 1, tile = 0;
  int local_dt = 0;  // AUSM_DT_LOCAL: each cell steps with its own CFL dt
  bool use_graphs = true;
  long long steps = 0;
  int since_refresh = 0;  // steps since the wave speed was last refreshed
  std::vector<ausm::GraphEntry> graphs;
  // all-FP64 solve (allocated on the first switch to FP64)
  int prec = 0;                                   // AUSM_PREC_FP32 / AUSM_PREC_FP64
  ausm::DevParamsT<double> hpar64{};              // free stream unrounded
  ausm::DevParamsT<double>* dpar64 = nullptr;
  double* q64[3] = {nullptr, nullptr, nullptr};   // Q^n, Q1, Q2 in FP64
  double* ws_part64 = nullptr;
  double* stage_buf64 = nullptr;                  // AoS staging, (N+3) x M x 4 doubles
  // compensated FP32 state (AUSM_PREC_FP32C): q[0] holds hi
  float* acc = nullptr;  // increment accumulator planes
  float* lo = nullptr;   // low part of the solution
  // limiter freezing: 2 while the record is in use, else 0; psi holds
  // 2 kPsiPerFace planes of psi_plane(N, P) values of the current precision
  // (sized for FP64, allocated on the first freeze)
  int freeze = 0;
  void* psi = nullptr;
};

namespace ausm {
int set_error(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
// AUSM_OK once the device has passed the exact arithmetic's self-check (run
// on the first call per device and process, with an allocation)
int device_arithmetic_ok(int device);
// the self-check's mismatch count on `device` in this process, -1 before it
// has run; storing a count; the error (or AUSM_OK) for a count
long long device_arithmetic_cached(int device);
void device_arithmetic_store(int device, long long bad);
int device_arithmetic_status(int device, long long bad);
size_t round_up(size_t x, size_t a);
// the FP32 parameter block of a free stream (inState formed in double,
// stored as FP32, as the reference's host does)
void make_freestream(DevParams& p, double rho, double pr, double vx, double vy, double gamma);
StageArgs stage_args(ausm_ctx* c, int ty, int stage);
StageArgs64 stage_args64(ausm_ctx* c, int ty, int stage);
}  // namespace ausm
