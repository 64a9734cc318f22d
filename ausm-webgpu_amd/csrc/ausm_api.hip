// C ABI of libausm: context management, geometry, grid generation, solver
// parameters, stepping (with hipGraph capture of step chunks and a tile
// autotuner) and read-back.  See include/ausm.h for the contract.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "aux_kernels.hpp"
#include "ausm.h"
#include "ctx.hpp"
#include "kernels.hpp"

using namespace ausm;

namespace ausm {

static thread_local std::string g_err;

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return set_error(AUSM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// The exact arithmetic divides with a refined hardware reciprocal that must
// be the correctly rounded 1/d (recip_rn, device_math.hpp) and takes square
// roots with a short form of the IEEE sequence (sqrt_rn); the check runs
// both over their whole ranges on the current device.
static int arithmetic_check(long long* mismatches) {
  unsigned long long* bad = nullptr;
  hipError_t e = hipMalloc(&bad, sizeof(unsigned long long));
  if (e != hipSuccess) return set_error(AUSM_ERR_NOMEM, "arithmetic check: %s", hipGetErrorString(e));
  unsigned long long h = 0;
  e = hipMemset(bad, 0, sizeof(h));
  if (e == hipSuccess) e = launch_arithmetic_check(bad, nullptr);
  if (e == hipSuccess) e = hipMemcpy(&h, bad, sizeof(h), hipMemcpyDeviceToHost);
  (void)hipFree(bad);
  if (e != hipSuccess) return hip_fail(e, "arithmetic check");
  *mismatches = (long long)h;
  return AUSM_OK;
}

// once per device and process: whether the check has run, and its mismatch count
static std::atomic<long long> g_arith_bad[64] = {};
static std::atomic<bool> g_arith_run[64] = {};

long long device_arithmetic_cached(int device) {
  if (device < 0 || device >= 64 || !g_arith_run[device].load()) return -1;
  return g_arith_bad[device].load();
}

void device_arithmetic_store(int device, long long bad) {
  if (device < 0 || device >= 64) return;
  g_arith_bad[device].store(bad);
  g_arith_run[device].store(true);
}

int device_arithmetic_status(int device, long long bad) {
  if (bad == 0) return AUSM_OK;
  return set_error(AUSM_ERR_HIP,
                   "device %d: the short division or square root of the exact arithmetic differs "
                   "from IEEE on %lld inputs (ausm_check_arithmetic)", device, bad);
}

int device_arithmetic_ok(int device) {
  long long bad = device_arithmetic_cached(device);
  if (bad < 0) {
    const int r = arithmetic_check(&bad);
    if (r != AUSM_OK) return r;
    device_arithmetic_store(device, bad);
  }
  return device_arithmetic_status(device, bad);
}

// Host-side parameter block -> device copy (the Uniforms upload,
// main.js:735).  Derived fields for the fast path are formed in double.
void fill_derived(DevParams& p) {
  const double g = p.gamma;
  p.gm1 = (float)(g - 1.0);
  p.ggm1 = (float)(g / (g - 1.0));
  p.minf_coef = (float)(((double)p.vx * p.vx + (double)p.vy * p.vy) / (g * (double)p.inP));
  p.astar_coef = (float)(2.0 * (g - 1.0) / (g + 1.0));
}

void make_freestream(DevParams& p, double rho, double pr, double vx, double vy, double gamma) {
  p.gamma = (float)gamma;
  p.inRho = (float)rho;
  p.inP = (float)pr;
  p.vx = (float)vx;
  p.vy = (float)vy;
  const double rhoE = pr / (gamma - 1.0) + 0.5 * (vx * vx + vy * vy) * rho;
  p.inState[0] = (float)rho;
  p.inState[1] = (float)(vx * rho);
  p.inState[2] = (float)(vy * rho);
  p.inState[3] = (float)rhoE;
  fill_derived(p);
}

// FP64 parameter block: the free stream unrounded; the solver settings the
// C ABI takes as float widened from the FP32 block.
void make_freestream64(DevParamsT<double>& p, double rho, double pr, double vx, double vy,
                       double gamma) {
  p.gamma = gamma;
  p.inRho = rho;
  p.inP = pr;
  p.vx = vx;
  p.vy = vy;
  p.inState[0] = rho;
  p.inState[1] = vx * rho;
  p.inState[2] = vy * rho;
  p.inState[3] = pr / (gamma - 1.0) + 0.5 * (vx * vx + vy * vy) * rho;
  p.gm1 = gamma - 1.0;
  p.ggm1 = gamma / (gamma - 1.0);
  p.minf_coef = (vx * vx + vy * vy) / (gamma * pr);
  p.astar_coef = 2.0 * (gamma - 1.0) / (gamma + 1.0);
}

static void widen_settings(ausm_ctx* c) {
  c->hpar64.maxdt = c->hpar.maxdt;
  c->hpar64.cfl = c->hpar.cfl;
  c->hpar64.Kp = c->hpar.Kp;
  c->hpar64.Ku = c->hpar.Ku;
}

static bool fp64(const ausm_ctx* c) { return c->prec == AUSM_PREC_FP64; }
static bool comp(const ausm_ctx* c) { return c->prec == AUSM_PREC_FP32C; }
static const double* gx64(const ausm_ctx* c) { return c->g64[c->g_cur]; }
static const double* gy64(const ausm_ctx* c) { return c->g64[c->g_cur + 1]; }

StageArgs64 stage_args64(ausm_ctx* c, int ty, int stage) {
  StageArgs64 a{};
  a.vx = gx64(c);
  a.vy = gy64(c);
  a.par = c->dpar64;
  a.ws_in = c->ws_part64;
  a.ws_out = c->ws_part64;
  a.res_out = c->res_part;
  a.maxws_cur = c->scal_d + 6;
  a.n_ws_in = c->n_ws_valid;
  a.M = c->M;
  a.N = c->N;
  a.P = c->P;
  a.PS = c->PS;
  a.stage = stage;
  a.limiter = c->limiter;
  a.tiles_x = stage_tiles_x(c->M);
  a.local_dt = c->local_dt;
  a.qin = c->q64[0];
  a.qn = c->q64[0];
  a.qout = c->q64[1];
  a.freeze = c->freeze;
  a.psi = (double*)c->psi;
  return a;
}

// One stage launch in the context's precision; in, qn and out index the
// three state buffers.  `norm`: stage 1 writes the residual-norm partials.
static hipError_t stage_launch(ausm_ctx* c, int ty, int stage, int in, int qn, int out, bool fresh,
                               bool refresh, bool norm) {
  if (fp64(c)) {
    StageArgs64 a = stage_args64(c, ty, stage);
    a.qin = c->q64[in];
    a.qn = c->q64[qn];
    a.qout = c->q64[out];
    a.fresh = fresh;
    a.refresh = refresh;
    if (!norm) a.res_out = nullptr;
    return launch_stage64(a, c->scheme, ty, c->stream);
  }
  StageArgs a = stage_args(c, ty, stage);
  a.qin = c->q[in];
  a.qn = c->q[qn];
  a.qout = c->q[out];
  a.fresh = fresh;
  a.refresh = refresh;
  if (!norm) a.res_out = nullptr;
  if (comp(c)) return launch_stage_comp(a, c->scheme, c->math, ty, c->stream);
  return launch_stage(a, c->scheme, c->math, ty, c->stream);
}

// ---- tile autotuner -------------------------------------------------------
// Rows per wavefront tile of the stage kernel, chosen per (M, N, scheme,
// math) by timing one stage on the context's own scratch buffer (Q2, which
// every step overwrites before reading), and remembered for the process.
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int, int, int>, int> g_tune;

static const int kTileChoices[] = {1, 2, 4, 8};

int autotune_rows(ausm_ctx* c) {
  const auto key = std::make_tuple(c->M, c->N, c->scheme, c->prec * 2 + c->math, c->limiter);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) return it->second;
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  // three rounds over the candidates, each a run of 30 launches; a
  // candidate's time is its best round (robust to clock and queue noise)
  float best_ms[4] = {1e30f, 1e30f, 1e30f, 1e30f};
  for (int round = 0; round < 3; ++round) {
    for (int k = 0; k < 4; ++k) {
      const int ty = kTileChoices[k];
      if (c->prec != AUSM_PREC_FP32 && ty > 4) continue;  // FP64 / compensated: 1, 2, 4 rows
      for (int w = 0; w < 3; ++w) (void)stage_launch(c, ty, 2, 0, 0, 2, false, false, false);
      (void)hipEventRecord(e0, c->stream);
      for (int r = 0; r < 30; ++r) (void)stage_launch(c, ty, 2, 0, 0, 2, false, false, false);
      (void)hipEventRecord(e1, c->stream);
      (void)hipEventSynchronize(e1);
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      best_ms[k] = std::min(best_ms[k], ms);
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  int best = kTileChoices[0];
  float best_t = best_ms[0];
  for (int k = 1; k < 4; ++k)
    if (best_ms[k] < best_t * 0.98f) {
      best_t = best_ms[k];
      best = kTileChoices[k];
    }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[key] = best;
  return best;
}

StageArgs stage_args(ausm_ctx* c, int ty, int stage) {
  StageArgs a{};
  a.vx = c->vx;
  a.vy = c->vy;
  a.par = c->dpar;
  a.ws_in = c->ws_part;
  a.ws_out = c->ws_part;
  a.res_out = c->res_part;
  a.maxws_cur = c->scal_f;
  a.n_ws_in = c->n_ws_valid;
  a.M = c->M;
  a.N = c->N;
  a.P = c->P;
  a.PS = c->PS;
  a.stage = stage;
  a.limiter = c->limiter;
  a.fresh = 0;
  a.refresh = 0;
  a.tiles_x = stage_tiles_x(c->M);
  a.local_dt = c->local_dt;
  a.acc = c->acc;
  a.lo = c->lo;
  a.qin = c->q[0];
  a.qn = c->q[0];
  a.qout = c->q[1];
  a.freeze = c->freeze;
  a.psi = (float*)c->psi;
  return a;
}

// Enqueue one TVD-RK3 step: three fused stage launches (the reference's 15
// passes, main.js:744-777).  `norm`: stage 1 also writes the partial sums of
// the density-residual norm (only the last step of a call needs them).
hipError_t enqueue_step(ausm_ctx* c, int ty, bool fresh, bool refresh, bool norm) {
  hipError_t e = hipSuccess;
  if (fresh && c->n_ws_valid > kWsInKernelMax) {
    // large grids: one launch reduces the maxima of the last refresh into
    // the stage's max wave speed, rather than every wavefront of stage 1
    // reading all of them (the max is exact, so dt is the same either way)
    e = fp64(c) ? launch_reduce_max(c->ws_part64, c->n_ws_valid, c->scal_d + 6, c->stream)
                : launch_reduce_max(c->ws_part, c->n_ws_valid, c->scal_f, c->stream);
    if (e != hipSuccess) return e;
    fresh = false;
  }
  e = stage_launch(c, ty, 1, 0, 0, 1, fresh, false, norm);
  if (e != hipSuccess) return e;
  e = stage_launch(c, ty, 2, 1, 0, 2, false, false, false);
  if (e != hipSuccess) return e;
  return stage_launch(c, ty, 3, 2, 0, 0, false, refresh, false);
}

// captured step chunks kept per context (full chunks and remainders of
// each tile shape and initial maxima count)
constexpr size_t kMaxGraphs = 16;

void drop_graphs(ausm_ctx* c) {
  // a launched graph may still be running on the stream
  if (!c->graphs.empty()) (void)hipStreamSynchronize(c->stream);
  for (auto& g : c->graphs) {
    hipGraphExecDestroy(g.exec);
    hipGraphDestroy(g.graph);
  }
  c->graphs.clear();
}

// a new state, grid, limiter or precision ends limiter freezing (the
// recorded psi belongs to the state it was taken from)
static void unfreeze(ausm_ctx* c) {
  if (!c->freeze) return;
  drop_graphs(c);  // captured stages carry the freeze mode
  c->freeze = 0;
}

}  // namespace ausm

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* ausm_last_error(void) { return g_err.c_str(); }
int ausm_version(void) { return AUSM_VERSION; }

#define CHECK_HIP(expr, what)                    \
  do {                                           \
    hipError_t _e = (expr);                      \
    if (_e != hipSuccess) return hip_fail(_e, what); \
  } while (0)

int ausm_ctx_create(int device, int M, int N, ausm_ctx** out) {
  if (!out) return set_error(AUSM_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (M < 4 || N < 4 || M > (1 << 15) || N > (1 << 15))
    return set_error(AUSM_ERR_ARG, "grid size %d x %d out of range (4..32768)", M, N);
  int ndev = 0;
  CHECK_HIP(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
  if (device < 0 || device >= ndev)
    return set_error(AUSM_ERR_ARG, "device %d not present (%d devices)", device, ndev);
  CHECK_HIP(hipSetDevice(device), "hipSetDevice");
  const int dr = device_arithmetic_ok(device);
  if (dr != AUSM_OK) return dr;
  ausm_ctx* c = new ausm_ctx();
  c->device = device;
  c->M = M;
  c->N = N;
  c->P = (int)round_up((size_t)M, 64);
  c->PS = (int)round_up((size_t)N * c->P, 64);
  int maxb = wave_speed_blocks(M, N);
  for (int ty : kTileChoices) maxb = std::max(maxb, stage_blocks(M, N, ty));
  c->max_parts = maxb;
  const size_t nv = (size_t)(N + 1) * c->P;
  bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
  for (int k = 0; k < 3 && ok; ++k) ok = hipMalloc(&c->q[k], sizeof(float) * 4 * c->PS) == hipSuccess;
  ok = ok && hipMalloc(&c->vx, sizeof(float) * nv) == hipSuccess;
  ok = ok && hipMalloc(&c->vy, sizeof(float) * nv) == hipSuccess;
  for (int k = 0; k < 4 && ok; ++k) ok = hipMalloc(&c->g64[k], sizeof(double) * nv) == hipSuccess;
  ok = ok && hipMalloc(&c->ws_part, sizeof(float) * maxb) == hipSuccess;
  ok = ok && hipMalloc(&c->res_part, sizeof(double) * (size_t)N * stage_tiles_x(M)) == hipSuccess;
  ok = ok && hipMalloc(&c->scal_f, sizeof(float) * 8) == hipSuccess;
  ok = ok && hipMalloc(&c->scal_d, sizeof(double) * 8) == hipSuccess;
  ok = ok && hipMalloc(&c->dpar, sizeof(DevParams)) == hipSuccess;
  ok = ok && hipMalloc(&c->dpar64, sizeof(DevParamsT<double>)) == hipSuccess;
  ok = ok && hipMalloc(&c->stage_buf, sizeof(float) * 4 * (size_t)M * (N + 3)) == hipSuccess;
  ok = ok && hipMalloc(&c->seg, sizeof(double) * (2 * (size_t)M + M)) == hipSuccess;
  ok = ok && hipHostMalloc(&c->pinned, 256) == hipSuccess;
  if (!ok) {
    ausm_ctx_destroy(c);
    return set_error(AUSM_ERR_NOMEM, "device allocation failed for a %d x %d context", M, N);
  }
  // defaults: the reference's start-up uniforms (main.js:874-887, util.js:76-99)
  make_freestream(c->hpar, 1.0, 1.0 / 1.4, 0.0, 0.0, 1.4);
  make_freestream64(c->hpar64, 1.0, 1.0 / 1.4, 0.0, 0.0, 1.4);
  c->hpar.maxdt = 1e-4f;
  c->hpar.cfl = 1.5f;
  c->hpar.Kp = 0.0f;
  c->hpar.Ku = 0.75f;
  widen_settings(c);
  c->scheme = SCHEME_SLAU2;
  c->limiter = LIM_VANALBADA;
  c->math = 0;
  c->refresh = 1;
  c->tile = 0;
  c->use_graphs = getenv("AUSM_NO_GRAPHS") == nullptr;
  (void)hipMemcpyAsync(c->dpar, &c->hpar, sizeof(DevParams), hipMemcpyHostToDevice, c->stream);
  (void)hipMemcpyAsync(c->dpar64, &c->hpar64, sizeof(DevParamsT<double>), hipMemcpyHostToDevice,
                       c->stream);
  (void)hipMemsetAsync(c->scal_f, 0, sizeof(float) * 8, c->stream);
  (void)hipMemsetAsync(c->scal_d, 0, sizeof(double) * 8, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    ausm_ctx_destroy(c);
    return set_error(AUSM_ERR_HIP, "context initialisation failed");
  }
  *out = c;
  return AUSM_OK;
}

int ausm_ctx_destroy(ausm_ctx* c) {
  if (!c) return AUSM_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  drop_graphs(c);
  for (int k = 0; k < 3; ++k) hipFree(c->q[k]);
  hipFree(c->vx);
  hipFree(c->vy);
  for (int k = 0; k < 4; ++k) hipFree(c->g64[k]);
  hipFree(c->ws_part);
  hipFree(c->res_part);
  hipFree(c->scal_f);
  hipFree(c->scal_d);
  hipFree(c->dpar);
  hipFree(c->dpar64);
  for (int k = 0; k < 3; ++k) hipFree(c->q64[k]);
  hipFree(c->ws_part64);
  hipFree(c->stage_buf64);
  hipFree(c->acc);
  hipFree(c->lo);
  hipFree(c->psi);
  hipFree(c->stage_buf);
  hipFree(c->seg);
  if (c->pinned) hipHostFree(c->pinned);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return AUSM_OK;
}

int ausm_ctx_stream(ausm_ctx* c, void** stream) {
  if (!c || !stream) return set_error(AUSM_ERR_ARG, "NULL argument");
  *stream = (void*)c->stream;
  return AUSM_OK;
}

int ausm_sync(ausm_ctx* c) {
  if (!c) return set_error(AUSM_ERR_ARG, "NULL context");
  CHECK_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  return AUSM_OK;
}

// ---- grid -------------------------------------------------------------------
static int finish_grid(ausm_ctx* c, const double* gx, const double* gy) {
  // captured FP64 step graphs hold the previous grid's vertex pointers
  // (g64[g_cur], which moves between the two ping-pong pairs)
  drop_graphs(c);
  unfreeze(c);
  const size_t nv = (size_t)(c->N + 1) * c->P;
  CHECK_HIP(launch_f64_to_f32(gx, c->vx, nv, c->stream), "grid convert");
  CHECK_HIP(launch_f64_to_f32(gy, c->vy, nv, c->stream), "grid convert");
  c->has_grid = true;
  c->has_state = false;
  return AUSM_OK;
}

int ausm_grid_from_boundary(ausm_ctx* c, const double* body, const double* outer, int iters,
                            int method, double omega) {
  if (!c || !body || !outer || iters < 0) return set_error(AUSM_ERR_ARG, "bad grid arguments");
  if (method != AUSM_GRID_JACOBI && method != AUSM_GRID_SOR)
    return set_error(AUSM_ERR_ARG, "unknown grid method %d", method);
  if (method == AUSM_GRID_SOR && (c->M % 2 != 0))
    return set_error(AUSM_ERR_ARG, "red-black SOR needs an even number of points around (M=%d)", c->M);
  hipSetDevice(c->device);
  const int M = c->M, N = c->N, P = c->P;
  // boundary rows into FP64 planes (writeTexture, main.js:665-670)
  std::vector<double> hx((size_t)(N + 1) * P, 0.0), hy((size_t)(N + 1) * P, 0.0);
  for (int i = 0; i < M; ++i) {
    hx[i] = body[2 * i];
    hy[i] = body[2 * i + 1];
    hx[(size_t)N * P + i] = outer[2 * i];
    hy[(size_t)N * P + i] = outer[2 * i + 1];
  }
  const size_t bytes = sizeof(double) * hx.size();
  double *gx = c->g64[0], *gy = c->g64[1], *hx2 = c->g64[2], *hy2 = c->g64[3];
  CHECK_HIP(hipMemcpyAsync(gx, hx.data(), bytes, hipMemcpyHostToDevice, c->stream), "grid upload");
  CHECK_HIP(hipMemcpyAsync(gy, hy.data(), bytes, hipMemcpyHostToDevice, c->stream), "grid upload");
  CHECK_HIP(launch_grid_interp(gx, gy, M, N, P, c->stream), "grid interpolation");
  if (method == AUSM_GRID_JACOBI) {
    // whole blocks of sweeps in one launch each, the rest one per launch
    const int B = grid_block_sweeps();
    for (int it = 0; it < iters;) {
      if (iters - it >= B) {
        CHECK_HIP(launch_grid_jacobi_block(gx, gy, hx2, hy2, M, N, P, c->stream), "grid Jacobi");
        it += B;
      } else {
        CHECK_HIP(launch_grid_jacobi(gx, gy, hx2, hy2, M, N, P, c->stream), "grid Jacobi");
        ++it;
      }
      std::swap(gx, hx2);
      std::swap(gy, hy2);
    }
  } else {
    // half-sweeps h = 0 .. 2 iters - 1 of colour h & 1, in blocks as above
    const int B = grid_block_sweeps();
    for (int h = 0; h < 2 * iters;) {
      if (2 * iters - h >= B) {
        CHECK_HIP(launch_grid_sor_block(gx, gy, hx2, hy2, M, N, P, h & 1, omega, c->stream),
                  "grid SOR");
        h += B;
      } else {
        CHECK_HIP(launch_grid_sor(gx, gy, hx2, hy2, M, N, P, h & 1, omega, c->stream), "grid SOR");
        ++h;
      }
      std::swap(gx, hx2);
      std::swap(gy, hy2);
    }
  }
  c->g_cur = (gx == c->g64[0]) ? 0 : 2;
  const int r = finish_grid(c, gx, gy);
  if (r != AUSM_OK) return r;
  CHECK_HIP(hipStreamSynchronize(c->stream), "grid generation");
  return AUSM_OK;
}

int ausm_grid_set_vertices(ausm_ctx* c, const float* xy) {
  if (!c || !xy) return set_error(AUSM_ERR_ARG, "NULL argument");
  hipSetDevice(c->device);
  const int M = c->M, N = c->N, P = c->P;
  std::vector<float> hx((size_t)(N + 1) * P, 0.0f), hy((size_t)(N + 1) * P, 0.0f);
  for (int j = 0; j <= N; ++j)
    for (int i = 0; i < M; ++i) {
      hx[(size_t)j * P + i] = xy[((size_t)j * M + i) * 2];
      hy[(size_t)j * P + i] = xy[((size_t)j * M + i) * 2 + 1];
    }
  CHECK_HIP(hipMemcpyAsync(c->vx, hx.data(), sizeof(float) * hx.size(), hipMemcpyHostToDevice, c->stream), "vertex upload");
  CHECK_HIP(hipMemcpyAsync(c->vy, hy.data(), sizeof(float) * hy.size(), hipMemcpyHostToDevice, c->stream), "vertex upload");
  // the FP64 solve uses the same vertices, widened
  std::vector<double> dx(hx.begin(), hx.end()), dy(hy.begin(), hy.end());
  CHECK_HIP(hipMemcpyAsync(c->g64[0], dx.data(), sizeof(double) * dx.size(), hipMemcpyHostToDevice, c->stream), "vertex upload");
  CHECK_HIP(hipMemcpyAsync(c->g64[1], dy.data(), sizeof(double) * dy.size(), hipMemcpyHostToDevice, c->stream), "vertex upload");
  CHECK_HIP(hipStreamSynchronize(c->stream), "vertex upload");
  drop_graphs(c);  // FP64 graphs may hold the other vertex pair
  unfreeze(c);
  c->has_grid = true;
  c->has_state = false;
  c->g_cur = 0;
  return AUSM_OK;
}

int ausm_grid_set_vertices64(ausm_ctx* c, const double* xy) {
  if (!c || !xy) return set_error(AUSM_ERR_ARG, "NULL argument");
  hipSetDevice(c->device);
  const int M = c->M, N = c->N, P = c->P;
  std::vector<double> hx((size_t)(N + 1) * P, 0.0), hy((size_t)(N + 1) * P, 0.0);
  for (int j = 0; j <= N; ++j)
    for (int i = 0; i < M; ++i) {
      hx[(size_t)j * P + i] = xy[((size_t)j * M + i) * 2];
      hy[(size_t)j * P + i] = xy[((size_t)j * M + i) * 2 + 1];
    }
  CHECK_HIP(hipMemcpyAsync(c->g64[0], hx.data(), sizeof(double) * hx.size(), hipMemcpyHostToDevice, c->stream), "vertex upload");
  CHECK_HIP(hipMemcpyAsync(c->g64[1], hy.data(), sizeof(double) * hy.size(), hipMemcpyHostToDevice, c->stream), "vertex upload");
  c->g_cur = 0;
  const int r = finish_grid(c, c->g64[0], c->g64[1]);
  if (r != AUSM_OK) return r;
  CHECK_HIP(hipStreamSynchronize(c->stream), "vertex upload");
  return AUSM_OK;
}

int ausm_grid_get_vertices(ausm_ctx* c, float* xy) {
  if (!c || !xy) return set_error(AUSM_ERR_ARG, "NULL argument");
  if (!c->has_grid) return set_error(AUSM_ERR_ORDER, "no grid");
  hipSetDevice(c->device);
  const int M = c->M, N = c->N, P = c->P;
  std::vector<float> hx((size_t)(N + 1) * P), hy((size_t)(N + 1) * P);
  CHECK_HIP(hipMemcpyAsync(hx.data(), c->vx, sizeof(float) * hx.size(), hipMemcpyDeviceToHost, c->stream), "vertex download");
  CHECK_HIP(hipMemcpyAsync(hy.data(), c->vy, sizeof(float) * hy.size(), hipMemcpyDeviceToHost, c->stream), "vertex download");
  CHECK_HIP(hipStreamSynchronize(c->stream), "vertex download");
  for (int j = 0; j <= N; ++j)
    for (int i = 0; i < M; ++i) {
      xy[((size_t)j * M + i) * 2] = hx[(size_t)j * P + i];
      xy[((size_t)j * M + i) * 2 + 1] = hy[(size_t)j * P + i];
    }
  return AUSM_OK;
}

int ausm_grid_get_vertices64(ausm_ctx* c, double* xy) {
  if (!c || !xy) return set_error(AUSM_ERR_ARG, "NULL argument");
  if (c->g_cur < 0) return set_error(AUSM_ERR_ORDER, "no grid");
  hipSetDevice(c->device);
  const int M = c->M, N = c->N, P = c->P;
  std::vector<double> hx((size_t)(N + 1) * P), hy((size_t)(N + 1) * P);
  CHECK_HIP(hipMemcpyAsync(hx.data(), c->g64[c->g_cur], sizeof(double) * hx.size(), hipMemcpyDeviceToHost, c->stream), "grid download");
  CHECK_HIP(hipMemcpyAsync(hy.data(), c->g64[c->g_cur + 1], sizeof(double) * hy.size(), hipMemcpyDeviceToHost, c->stream), "grid download");
  CHECK_HIP(hipStreamSynchronize(c->stream), "grid download");
  for (int j = 0; j <= N; ++j)
    for (int i = 0; i < M; ++i) {
      xy[((size_t)j * M + i) * 2] = hx[(size_t)j * P + i];
      xy[((size_t)j * M + i) * 2 + 1] = hy[(size_t)j * P + i];
    }
  return AUSM_OK;
}

int ausm_grid_get_metrics(ausm_ctx* c, float* area, float* flen) {
  if (!c || !area || !flen) return set_error(AUSM_ERR_ARG, "NULL argument");
  if (!c->has_grid) return set_error(AUSM_ERR_ORDER, "no grid");
  hipSetDevice(c->device);
  const int M = c->M, N = c->N;
  float* da = c->stage_buf;  // M*N
  float* df = c->q[2];       // M*N*4: the scratch stage buffer Q2 is dead between steps
  CHECK_HIP(launch_metrics(c->vx, c->vy, M, N, c->P, da, df, c->stream), "metrics");
  CHECK_HIP(hipMemcpyAsync(area, da, sizeof(float) * M * N, hipMemcpyDeviceToHost, c->stream), "metrics download");
  CHECK_HIP(hipMemcpyAsync(flen, df, sizeof(float) * 4 * M * N, hipMemcpyDeviceToHost, c->stream), "metrics download");
  CHECK_HIP(hipStreamSynchronize(c->stream), "metrics");
  return AUSM_OK;
}

// ---- parameters ---------------------------------------------------------------
static int upload_params(ausm_ctx* c) {
  widen_settings(c);
  CHECK_HIP(hipMemcpyAsync(c->dpar, &c->hpar, sizeof(DevParams), hipMemcpyHostToDevice, c->stream), "parameter upload");
  CHECK_HIP(hipMemcpyAsync(c->dpar64, &c->hpar64, sizeof(DevParamsT<double>), hipMemcpyHostToDevice, c->stream), "parameter upload");
  CHECK_HIP(hipStreamSynchronize(c->stream), "parameter upload");
  return AUSM_OK;
}

int ausm_set_freestream(ausm_ctx* c, double rho, double p, double vx, double vy, double gamma) {
  if (!c) return set_error(AUSM_ERR_ARG, "NULL context");
  if (!(rho > 0) || !(p > 0) || !(gamma > 1) || !std::isfinite(vx) || !std::isfinite(vy))
    return set_error(AUSM_ERR_ARG, "bad free stream (rho=%g p=%g gamma=%g)", rho, p, gamma);
  hipSetDevice(c->device);
  make_freestream(c->hpar, rho, p, vx, vy, gamma);
  make_freestream64(c->hpar64, rho, p, vx, vy, gamma);
  return upload_params(c);
}

int ausm_set_scheme(ausm_ctx* c, int scheme, int limiter, float K_p, float K_u) {
  if (!c) return set_error(AUSM_ERR_ARG, "NULL context");
  if (scheme < 0 || scheme > 2) return set_error(AUSM_ERR_ARG, "unknown scheme %d", scheme);
  if (limiter < 0 || limiter > 4) return set_error(AUSM_ERR_ARG, "unknown limiter %d", limiter);
  hipSetDevice(c->device);
  if (scheme != c->scheme || limiter != c->limiter) {
    drop_graphs(c);
    unfreeze(c);
  }
  c->scheme = scheme;
  c->limiter = limiter;
  c->hpar.Kp = K_p;
  c->hpar.Ku = K_u;
  return upload_params(c);
}

int ausm_set_cfl(ausm_ctx* c, float cfl, float maxdt, int refresh) {
  if (!c) return set_error(AUSM_ERR_ARG, "NULL context");
  if (!(cfl > 0) || !(maxdt >= 0) || refresh < 1)
    return set_error(AUSM_ERR_ARG, "bad CFL settings (cfl=%g maxdt=%g refresh=%d)", cfl, maxdt, refresh);
  hipSetDevice(c->device);
  if (refresh != c->refresh) drop_graphs(c);
  c->hpar.cfl = cfl;
  c->hpar.maxdt = maxdt;
  c->refresh = refresh;
  return upload_params(c);
}

int ausm_set_math(ausm_ctx* c, int math) {
  if (!c) return set_error(AUSM_ERR_ARG, "NULL context");
  if (math != AUSM_MATH_EXACT && math != AUSM_MATH_FAST)
    return set_error(AUSM_ERR_ARG, "unknown math mode %d", math);
  if (math != c->math) drop_graphs(c);
  c->math = math;
  return AUSM_OK;
}

int ausm_set_dt_mode(ausm_ctx* c, int mode) {
  if (!c) return set_error(AUSM_ERR_ARG, "NULL context");
  if (mode != AUSM_DT_GLOBAL && mode != AUSM_DT_LOCAL)
    return set_error(AUSM_ERR_ARG, "unknown dt mode %d", mode);
  if (mode != c->local_dt) drop_graphs(c);
  c->local_dt = mode;
  return AUSM_OK;
}

int ausm_set_tile(ausm_ctx* c, int rows) {
  if (!c) return set_error(AUSM_ERR_ARG, "NULL context");
  if (rows != 0 && rows != 1 && rows != 2 && rows != 4 && rows != 8)
    return set_error(AUSM_ERR_ARG, "rows per wavefront must be 0 (auto), 1, 2, 4 or 8");
  if (rows != c->tile) drop_graphs(c);
  c->tile = rows;
  return AUSM_OK;
}

// ---- limiter freezing ----------------------------------------------------------
// Records psi = phi(r)/r of every face and component from the current state
// (a residual pass of the general stage kernel) and makes every later stage
// reconstruct with it, until on = 0 or a new state, grid, limiter or
// precision.  A build extension with no counterpart in the reference.
static int current_rows(ausm_ctx* c);

int ausm_freeze_limiter(ausm_ctx* c, int on) {
  if (!c) return set_error(AUSM_ERR_ARG, "NULL context");
  if (on != 0 && on != 1) return set_error(AUSM_ERR_ARG, "freeze flag must be 0 or 1");
  hipSetDevice(c->device);
  if (!on) {
    unfreeze(c);
    return AUSM_OK;
  }
  if (!c->has_state) return set_error(AUSM_ERR_ORDER, "no state to record the limiter from");
  if (c->limiter != LIM_VANALBADA && c->limiter != LIM_VANALBADA_CLAMPED)
    return set_error(AUSM_ERR_ARG, "limiter freezing needs a van Albada limiter (limiter %d)", c->limiter);
  if (!c->psi && hipMalloc(&c->psi, sizeof(double) * 2 * kPsiPerFace * psi_plane(c->N, c->P)) != hipSuccess) {
    c->psi = nullptr;
    return set_error(AUSM_ERR_NOMEM, "device allocation failed for the limiter record");
  }
  const int ty = current_rows(c);  // tuned (if at all) before recording
  drop_graphs(c);
  c->freeze = 1;
  const hipError_t e = stage_launch(c, ty, 0, 0, 0, 2, false, false, false);
  c->freeze = e == hipSuccess ? 2 : 0;
  CHECK_HIP(e, "limiter record");
  CHECK_HIP(hipStreamSynchronize(c->stream), "limiter record");
  return AUSM_OK;
}

int ausm_limiter_frozen(ausm_ctx* c, int* on) {
  if (!c || !on) return set_error(AUSM_ERR_ARG, "NULL argument");
  *on = c->freeze ? 1 : 0;
  return AUSM_OK;
}

// ---- state --------------------------------------------------------------------
static int refresh_ws(ausm_ctx* c) {
  // wave speeds of the current solution -> per-block maxima (the post pass
  // of prepareState, main.js:649-652)
  if (fp64(c))
    CHECK_HIP(launch_wave_speed(c->q64[0], gx64(c), gy64(c), c->dpar64, c->M, c->N, c->P, c->PS,
                                (double*)nullptr, c->ws_part64, c->stream),
              "wave speeds");
  else
    CHECK_HIP(launch_wave_speed(c->q[0], c->vx, c->vy, c->dpar, c->M, c->N, c->P, c->PS,
                                (float*)nullptr, c->ws_part, c->stream),
              "wave speeds");
  c->n_ws_valid = wave_speed_blocks(c->M, c->N);
  c->ws_fresh = true;
  return AUSM_OK;
}

int ausm_init_state(ausm_ctx* c, int mode) {
  if (!c) return set_error(AUSM_ERR_ARG, "NULL context");
  if (!c->has_grid) return set_error(AUSM_ERR_ORDER, "set a grid before the state");
  if (mode != AUSM_INIT_REST && mode != AUSM_INIT_FREESTREAM)
    return set_error(AUSM_ERR_ARG, "unknown init mode %d", mode);
  hipSetDevice(c->device);
  unfreeze(c);
  if (fp64(c))
    CHECK_HIP(launch_fill_state(c->q64[0], c->dpar64, c->M, c->N, c->P, c->PS, mode, c->stream), "init state");
  else
    CHECK_HIP(launch_fill_state(c->q[0], c->dpar, c->M, c->N, c->P, c->PS, mode, c->stream), "init state");
  if (comp(c)) CHECK_HIP(hipMemsetAsync(c->lo, 0, sizeof(float) * 4 * (size_t)c->PS, c->stream), "init state");
  int r = refresh_ws(c);
  if (r != AUSM_OK) return r;
  CHECK_HIP(hipStreamSynchronize(c->stream), "init state");
  c->has_state = true;
  c->steps = 0;
  c->since_refresh = 0;
  return AUSM_OK;
}

int ausm_set_state(ausm_ctx* c, const float* q) {
  if (!c || !q) return set_error(AUSM_ERR_ARG, "NULL argument");
  if (!c->has_grid) return set_error(AUSM_ERR_ORDER, "set a grid before the state");
  hipSetDevice(c->device);
  unfreeze(c);
  const size_t n = 4 * (size_t)c->M * (c->N + 3);
  CHECK_HIP(hipMemcpyAsync(c->stage_buf, q, sizeof(float) * n, hipMemcpyHostToDevice, c->stream), "state upload");
  if (fp64(c)) {
    CHECK_HIP(launch_aos_to_soa(c->stage_buf, c->q[1], c->M, c->N, c->P, c->PS, c->stream), "state upload");
    CHECK_HIP(launch_convert_planes(c->q[1], c->q64[0], c->M, c->N, c->P, c->PS, c->stream), "state upload");
  } else {
    CHECK_HIP(launch_aos_to_soa(c->stage_buf, c->q[0], c->M, c->N, c->P, c->PS, c->stream), "state upload");
    if (comp(c)) CHECK_HIP(hipMemsetAsync(c->lo, 0, sizeof(float) * 4 * (size_t)c->PS, c->stream), "state upload");
  }
  int r = refresh_ws(c);
  if (r != AUSM_OK) return r;
  CHECK_HIP(hipStreamSynchronize(c->stream), "state upload");
  c->has_state = true;
  c->steps = 0;
  c->since_refresh = 0;
  return AUSM_OK;
}

int ausm_get_state(ausm_ctx* c, float* q) {
  if (!c || !q) return set_error(AUSM_ERR_ARG, "NULL argument");
  if (!c->has_state) return set_error(AUSM_ERR_ORDER, "no state");
  hipSetDevice(c->device);
  const size_t n = 4 * (size_t)c->M * (c->N + 3);
  if (fp64(c)) {
    CHECK_HIP(launch_soa_to_aos(c->q64[0], gx64(c), gy64(c), c->dpar64, c->M, c->N, c->P, c->PS, c->stage_buf64, c->stream), "state download");
    CHECK_HIP(launch_f64_to_f32(c->stage_buf64, c->stage_buf, n, c->stream), "state download");
  } else {
    CHECK_HIP(launch_soa_to_aos(c->q[0], c->vx, c->vy, c->dpar, c->M, c->N, c->P, c->PS, c->stage_buf, c->stream), "state download");
  }
  CHECK_HIP(hipMemcpyAsync(q, c->stage_buf, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream), "state download");
  CHECK_HIP(hipStreamSynchronize(c->stream), "state download");
  return AUSM_OK;
}

static int current_rows(ausm_ctx* c) {
  const int t = c->tile ? c->tile : autotune_rows(c);
  return c->prec != AUSM_PREC_FP32 ? std::min(t, 4) : t;  // FP64 / compensated: 1, 2, 4 rows
}

static int ensure_fp64_buffers(ausm_ctx* c) {
  if (c->q64[0]) return AUSM_OK;
  const size_t nq = sizeof(double) * 4 * (size_t)c->PS;
  bool ok = true;
  for (int k = 0; k < 3 && ok; ++k) ok = hipMalloc(&c->q64[k], nq) == hipSuccess;
  ok = ok && hipMalloc(&c->ws_part64, sizeof(double) * c->max_parts) == hipSuccess;
  ok = ok && hipMalloc(&c->stage_buf64, sizeof(double) * 4 * (size_t)c->M * (c->N + 3)) == hipSuccess;
  if (!ok) {
    for (int k = 0; k < 3; ++k) {
      hipFree(c->q64[k]);
      c->q64[k] = nullptr;
    }
    hipFree(c->ws_part64);
    hipFree(c->stage_buf64);
    c->ws_part64 = nullptr;
    c->stage_buf64 = nullptr;
    return set_error(AUSM_ERR_NOMEM, "device allocation failed for the FP64 state");
  }
  return AUSM_OK;
}

static int ensure_comp_buffers(ausm_ctx* c) {
  if (c->acc) return AUSM_OK;
  const size_t nq = sizeof(float) * 4 * (size_t)c->PS;
  if (hipMalloc(&c->acc, nq) != hipSuccess || hipMalloc(&c->lo, nq) != hipSuccess) {
    hipFree(c->acc);
    hipFree(c->lo);
    c->acc = c->lo = nullptr;
    return set_error(AUSM_ERR_NOMEM, "device allocation failed for the compensated state");
  }
  return AUSM_OK;
}

int ausm_set_precision(ausm_ctx* c, int precision) {
  if (!c) return set_error(AUSM_ERR_ARG, "NULL context");
  if (precision != AUSM_PREC_FP32 && precision != AUSM_PREC_FP64 && precision != AUSM_PREC_FP32C)
    return set_error(AUSM_ERR_ARG, "unknown precision %d", precision);
  if (precision == c->prec) return AUSM_OK;
  hipSetDevice(c->device);
  int r = AUSM_OK;
  if (precision == AUSM_PREC_FP64) r = ensure_fp64_buffers(c);
  if (r == AUSM_OK && precision == AUSM_PREC_FP32C) r = ensure_comp_buffers(c);
  if (r != AUSM_OK) return r;
  drop_graphs(c);
  unfreeze(c);
  // carry the current solution over to the new precision
  if (c->has_state) {
    const int M = c->M, N = c->N, P = c->P, PS = c->PS;
    const int from = c->prec, to = precision;
    if (to == AUSM_PREC_FP32C)
      CHECK_HIP(hipMemsetAsync(c->lo, 0, sizeof(float) * 4 * (size_t)PS, c->stream), "precision switch");
    if (from == AUSM_PREC_FP64 && to == AUSM_PREC_FP32C)
      CHECK_HIP(launch_f64_to_comp(c->q64[0], c->q[0], c->lo, M, N, P, PS, c->stream), "precision switch");
    else if (from == AUSM_PREC_FP64)
      CHECK_HIP(launch_convert_planes(c->q64[0], c->q[0], M, N, P, PS, c->stream), "precision switch");
    else if (to == AUSM_PREC_FP64 && from == AUSM_PREC_FP32C)
      CHECK_HIP(launch_comp_to_f64(c->q[0], c->lo, c->q64[0], M, N, P, PS, c->stream), "precision switch");
    else if (to == AUSM_PREC_FP64)
      CHECK_HIP(launch_convert_planes(c->q[0], c->q64[0], M, N, P, PS, c->stream), "precision switch");
    // FP32 <-> FP32C: hi is the FP32 state (round(hi + lo) = hi)
  }
  c->prec = precision;
  if (c->has_state) {
    r = refresh_ws(c);
    if (r != AUSM_OK) return r;
    c->since_refresh = 0;
  }
  CHECK_HIP(hipStreamSynchronize(c->stream), "precision switch");
  return AUSM_OK;
}

int ausm_set_state64(ausm_ctx* c, const double* q) {
  if (!c || !q) return set_error(AUSM_ERR_ARG, "NULL argument");
  if (!fp64(c) && !comp(c))
    return set_error(AUSM_ERR_ORDER, "FP64 state needs FP64 or compensated precision");
  if (!c->has_grid) return set_error(AUSM_ERR_ORDER, "set a grid before the state");
  hipSetDevice(c->device);
  unfreeze(c);
  const int rb = ensure_fp64_buffers(c);
  if (rb != AUSM_OK) return rb;
  const size_t n = 4 * (size_t)c->M * (c->N + 3);
  CHECK_HIP(hipMemcpyAsync(c->stage_buf64, q, sizeof(double) * n, hipMemcpyHostToDevice, c->stream), "state upload");
  CHECK_HIP(launch_aos_to_soa(c->stage_buf64, c->q64[0], c->M, c->N, c->P, c->PS, c->stream), "state upload");
  // compensated: split into hi + lo
  if (comp(c)) CHECK_HIP(launch_f64_to_comp(c->q64[0], c->q[0], c->lo, c->M, c->N, c->P, c->PS, c->stream), "state upload");
  const int r = refresh_ws(c);
  if (r != AUSM_OK) return r;
  CHECK_HIP(hipStreamSynchronize(c->stream), "state upload");
  c->has_state = true;
  c->steps = 0;
  c->since_refresh = 0;
  return AUSM_OK;
}

int ausm_get_state64(ausm_ctx* c, double* q) {
  if (!c || !q) return set_error(AUSM_ERR_ARG, "NULL argument");
  if (!fp64(c) && !comp(c))
    return set_error(AUSM_ERR_ORDER, "FP64 state needs FP64 or compensated precision");
  if (!c->has_state) return set_error(AUSM_ERR_ORDER, "no state");
  hipSetDevice(c->device);
  const int rb = ensure_fp64_buffers(c);
  if (rb != AUSM_OK) return rb;
  const size_t n = 4 * (size_t)c->M * (c->N + 3);
  // compensated: hi + lo (exact in FP64); the ghost rows are formed in FP64
  if (comp(c)) CHECK_HIP(launch_comp_to_f64(c->q[0], c->lo, c->q64[0], c->M, c->N, c->P, c->PS, c->stream), "state download");
  CHECK_HIP(launch_soa_to_aos(c->q64[0], gx64(c), gy64(c), c->dpar64, c->M, c->N, c->P, c->PS, c->stage_buf64, c->stream), "state download");
  CHECK_HIP(hipMemcpyAsync(q, c->stage_buf64, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream), "state download");
  CHECK_HIP(hipStreamSynchronize(c->stream), "state download");
  return AUSM_OK;
}

namespace ausm {

// The graph of `len` steps (len a multiple of the refresh cadence K) run from
// a state at a refresh boundary whose last refresh left n_ws_in per-workgroup
// maxima: found in the context's cache, or captured and instantiated (nothing
// runs).  Within the chunk step s is fresh iff s % K == 0.
static int graph_for(ausm_ctx* c, int ty, int len, int n_ws_in, GraphEntry** out) {
  for (auto& g : c->graphs)
    if (g.ty == ty && g.steps == len && g.n_ws_in == n_ws_in) {
      *out = &g;
      return AUSM_OK;
    }
  if (c->graphs.size() >= kMaxGraphs) {
    (void)hipStreamSynchronize(c->stream);
    // oldest first: the full-length chunks are captured first and
    // recaptured at once if they are ever needed again
    hipGraphExecDestroy(c->graphs.front().exec);
    hipGraphDestroy(c->graphs.front().graph);
    c->graphs.erase(c->graphs.begin());
  }
  const int K = c->refresh;
  GraphEntry g{};
  g.ty = ty;
  g.steps = len;
  g.n_ws_in = n_ws_in;
  const int save_valid = c->n_ws_valid;
  c->n_ws_valid = n_ws_in;
  CHECK_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "graph capture");
  hipError_t e = hipSuccess;
  for (int s = 0; s < len && e == hipSuccess; ++s) {
    const bool fresh = (s % K) == 0;
    const bool refresh = ((s + 1) % K) == 0;
    e = enqueue_step(c, ty, fresh, refresh, s == len - 1);
    if (refresh) c->n_ws_valid = stage_blocks(c->M, c->N, ty);
  }
  hipGraph_t graph = nullptr;
  hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
  c->n_ws_valid = save_valid;
  if (e != hipSuccess || e2 != hipSuccess) {
    if (graph) hipGraphDestroy(graph);
    return e != hipSuccess ? hip_fail(e, "graph capture (launch)") : hip_fail(e2, "graph capture");
  }
  g.graph = graph;
  CHECK_HIP(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0), "graph instantiate");
  // upload the executable graph to the device now, not at its first launch
  const hipError_t eu = hipGraphUpload(g.exec, c->stream);
  if (eu != hipSuccess) {
    hipGraphExecDestroy(g.exec);
    hipGraphDestroy(graph);
    return hip_fail(eu, "graph upload");
  }
  c->graphs.push_back(g);
  *out = &c->graphs.back();
  return AUSM_OK;
}

// graph chunk: a multiple of the dt-refresh cadence, about 32 steps; a
// shorter remainder of whole refresh periods is a graph of its own, so only
// the steps of a partial period run as single launches
static int chunk_steps(int K) { return K * std::max(1, 32 / K); }

}  // namespace ausm

int ausm_prepare_steps(ausm_ctx* c, int nsteps) {
  if (!c) return set_error(AUSM_ERR_ARG, "NULL context");
  if (nsteps < 0) return set_error(AUSM_ERR_ARG, "negative step count");
  if (!c->has_state) return set_error(AUSM_ERR_ORDER, "initialise the state before stepping");
  if (!c->use_graphs) return AUSM_OK;
  hipSetDevice(c->device);
  const int ty = current_rows(c);
  const int K = c->refresh;
  const int G = chunk_steps(K);
  const int steady = stage_blocks(c->M, c->N, ty);
  // the chunk sequence ausm_step would run from the current counters, and
  // from a refresh boundary left by an earlier step (what a call after
  // other steps starts from)
  for (int from = 0; from < 2; ++from) {
    if (from == 1 && c->since_refresh == 0 && c->ws_fresh && c->n_ws_valid == steady) break;
    int since = from ? 0 : c->since_refresh, valid = from ? steady : c->n_ws_valid;
    bool ws_fresh = from ? true : c->ws_fresh;
    int left = nsteps;
    while (left > 0) {
      if (since == 0 && ws_fresh && left >= K) {
        const int len = left >= G ? G : left / K * K;
        GraphEntry* ge = nullptr;
        const int r = graph_for(c, ty, len, valid, &ge);
        if (r != AUSM_OK) return r;
        left -= len;
        valid = steady;
      } else {
        const bool refresh = since + 1 >= K;
        ws_fresh = refresh;
        since = refresh ? 0 : since + 1;
        if (refresh) valid = steady;
        left -= 1;
      }
    }
  }
  return AUSM_OK;
}

int ausm_step(ausm_ctx* c, int nsteps, ausm_step_info* info) {
  if (!c) return set_error(AUSM_ERR_ARG, "NULL context");
  if (nsteps < 0) return set_error(AUSM_ERR_ARG, "negative step count");
  if (!c->has_state) return set_error(AUSM_ERR_ORDER, "initialise the state before stepping");
  hipSetDevice(c->device);
  const int ty = current_rows(c);
  const int K = c->refresh;
  const int G = chunk_steps(K);
  int left = nsteps;
  while (left > 0) {
    const bool aligned = c->since_refresh == 0 && c->ws_fresh;
    if (aligned && left >= K && c->use_graphs) {
      const int len = left >= G ? G : left / K * K;
      GraphEntry* ge = nullptr;
      const int r = graph_for(c, ty, len, c->n_ws_valid, &ge);
      if (r != AUSM_OK) return r;
      CHECK_HIP(hipGraphLaunch(ge->exec, c->stream), "graph launch");
      c->steps += len;
      left -= len;
      c->n_ws_valid = stage_blocks(c->M, c->N, ty);
      c->ws_fresh = true;
    } else {
      const bool fresh = c->ws_fresh;
      const bool refresh = c->since_refresh + 1 >= K;
      CHECK_HIP(enqueue_step(c, ty, fresh, refresh, left == 1), "stage launch");
      c->ws_fresh = refresh;
      c->since_refresh = refresh ? 0 : c->since_refresh + 1;
      if (refresh) c->n_ws_valid = stage_blocks(c->M, c->N, ty);
      c->steps += 1;
      left -= 1;
    }
  }
  if (info) {
    // density-residual norm of the last step's stage 1 and the current max
    // wave speed; one small read-back
    CHECK_HIP(launch_reduce_sum_f64(c->res_part, c->N * stage_tiles_x(c->M), c->scal_d, c->stream), "residual norm");
    double l2, cur, nxt, dt;
    if (fp64(c)) {
      CHECK_HIP(launch_reduce_max(c->ws_part64, c->n_ws_valid, c->scal_d + 7, c->stream), "max wave speed");
      CHECK_HIP(hipMemcpyAsync(c->pinned, c->scal_d, sizeof(double), hipMemcpyDeviceToHost, c->stream), "info");
      CHECK_HIP(hipMemcpyAsync((char*)c->pinned + 8, c->scal_d + 6, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream), "info");
      CHECK_HIP(hipStreamSynchronize(c->stream), "step");
      const double* h = (const double*)c->pinned;
      l2 = h[0];
      cur = h[1];
      nxt = h[2];
      dt = std::fmin(c->hpar64.maxdt, c->hpar64.cfl / cur);
    } else {
      CHECK_HIP(launch_reduce_max(c->ws_part, c->n_ws_valid, c->scal_f + 1, c->stream), "max wave speed");
      CHECK_HIP(hipMemcpyAsync(c->pinned, c->scal_d, sizeof(double), hipMemcpyDeviceToHost, c->stream), "info");
      CHECK_HIP(hipMemcpyAsync((char*)c->pinned + 8, c->scal_f, 2 * sizeof(float), hipMemcpyDeviceToHost, c->stream), "info");
      CHECK_HIP(hipStreamSynchronize(c->stream), "step");
      l2 = *(double*)c->pinned;
      const float* h = (const float*)((char*)c->pinned + 8);
      cur = h[0];
      nxt = h[1];
      dt = std::fmin(c->hpar.maxdt, c->hpar.cfl / h[0]);
    }
    info->res_l2 = nsteps > 0 ? l2 : -1.0;
    info->max_wave_speed = (float)cur;
    info->dt = (float)dt;
    info->nonfinite = !(std::isfinite(cur) && std::isfinite(nxt)) || !std::isfinite(l2);
    info->steps_done = c->steps;
    if (info->nonfinite) return set_error(AUSM_ERR_NONFINITE, "non-finite solution after %lld steps", c->steps);
  } else {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "step");
  }
  return AUSM_OK;
}

// stage 0 writes L(Q^n) into the scratch buffer Q2 (every step rewrites it
// before reading it)
static int residual_to_staging(ausm_ctx* c) {
  const int ty = current_rows(c);
  CHECK_HIP(stage_launch(c, ty, 0, 0, 0, 2, false, false, false), "residual");
  if (fp64(c))
    CHECK_HIP(launch_planes_to_aos(c->q64[2], c->M, c->N, c->P, c->PS, c->stage_buf64, c->stream), "residual");
  else
    CHECK_HIP(launch_planes_to_aos(c->q[2], c->M, c->N, c->P, c->PS, c->stage_buf, c->stream), "residual");
  return AUSM_OK;
}

int ausm_eval_residual(ausm_ctx* c, float* res) {
  if (!c || !res) return set_error(AUSM_ERR_ARG, "NULL argument");
  if (!c->has_state) return set_error(AUSM_ERR_ORDER, "no state");
  hipSetDevice(c->device);
  const size_t n = 4 * (size_t)c->M * c->N;
  const int r = residual_to_staging(c);
  if (r != AUSM_OK) return r;
  if (fp64(c)) CHECK_HIP(launch_f64_to_f32(c->stage_buf64, c->stage_buf, n, c->stream), "residual");
  CHECK_HIP(hipMemcpyAsync(res, c->stage_buf, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream), "residual download");
  CHECK_HIP(hipStreamSynchronize(c->stream), "residual");
  return AUSM_OK;
}

int ausm_eval_residual64(ausm_ctx* c, double* res) {
  if (!c || !res) return set_error(AUSM_ERR_ARG, "NULL argument");
  if (!fp64(c)) return set_error(AUSM_ERR_ORDER, "FP64 residual needs FP64 precision");
  if (!c->has_state) return set_error(AUSM_ERR_ORDER, "no state");
  hipSetDevice(c->device);
  const int r = residual_to_staging(c);
  if (r != AUSM_OK) return r;
  CHECK_HIP(hipMemcpyAsync(res, c->stage_buf64, sizeof(double) * 4 * (size_t)c->M * c->N, hipMemcpyDeviceToHost, c->stream), "residual download");
  CHECK_HIP(hipStreamSynchronize(c->stream), "residual");
  return AUSM_OK;
}

int ausm_wave_speeds(ausm_ctx* c, float* ws, float* max_ws) {
  if (!c) return set_error(AUSM_ERR_ARG, "NULL context");
  if (!c->has_state) return set_error(AUSM_ERR_ORDER, "no state");
  hipSetDevice(c->device);
  const size_t mn = (size_t)c->M * c->N;
  float* dws = c->stage_buf;
  const int nb = wave_speed_blocks(c->M, c->N);
  if (fp64(c)) {
    // FP64 speeds and maximum, returned rounded to FP32
    double* ws64 = c->stage_buf64;
    double* part = c->stage_buf64 + mn;
    CHECK_HIP(launch_wave_speed(c->q64[0], gx64(c), gy64(c), c->dpar64, c->M, c->N, c->P, c->PS, ws64, part, c->stream), "wave speeds");
    CHECK_HIP(launch_reduce_max(part, nb, c->scal_d + 7, c->stream), "max wave speed");
    CHECK_HIP(launch_f64_to_f32(ws64, dws, mn, c->stream), "wave speeds");
    CHECK_HIP(launch_f64_to_f32(c->scal_d + 7, c->scal_f + 2, 1, c->stream), "wave speeds");
  } else {
    float* part = c->stage_buf + mn;
    CHECK_HIP(launch_wave_speed(c->q[0], c->vx, c->vy, c->dpar, c->M, c->N, c->P, c->PS, dws, part, c->stream), "wave speeds");
    CHECK_HIP(launch_reduce_max(part, nb, c->scal_f + 2, c->stream), "max wave speed");
  }
  if (ws) CHECK_HIP(hipMemcpyAsync(ws, dws, sizeof(float) * mn, hipMemcpyDeviceToHost, c->stream), "wave speed download");
  CHECK_HIP(hipMemcpyAsync(c->pinned, c->scal_f + 2, sizeof(float), hipMemcpyDeviceToHost, c->stream), "wave speed download");
  CHECK_HIP(hipStreamSynchronize(c->stream), "wave speeds");
  if (max_ws) *max_ws = *(float*)c->pinned;
  return AUSM_OK;
}

int ausm_forces(ausm_ctx* c, double ref_length, double* out, double* cp) {
  if (!c || !out) return set_error(AUSM_ERR_ARG, "NULL argument");
  if (!(ref_length > 0)) return set_error(AUSM_ERR_ARG, "reference length must be positive");
  if (!c->has_state) return set_error(AUSM_ERR_ORDER, "no state");
  hipSetDevice(c->device);
  double* dcp = c->seg + 2 * (size_t)c->M;
  if (fp64(c))
    CHECK_HIP(launch_forces(c->q64[0], gx64(c), gy64(c), c->dpar64, c->M, c->PS, ref_length, c->seg, c->scal_d + 2, dcp, c->stream), "forces");
  else
    CHECK_HIP(launch_forces(c->q[0], c->vx, c->vy, c->dpar, c->M, c->PS, ref_length, c->seg, c->scal_d + 2, dcp, c->stream), "forces");
  CHECK_HIP(hipMemcpyAsync(out, c->scal_d + 2, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream), "forces download");
  if (cp) CHECK_HIP(hipMemcpyAsync(cp, dcp, sizeof(double) * c->M, hipMemcpyDeviceToHost, c->stream), "cp download");
  CHECK_HIP(hipStreamSynchronize(c->stream), "forces");
  return AUSM_OK;
}

int ausm_derived_field(ausm_ctx* c, int kind, float* out, float* freestream_value) {
  if (!c || !out) return set_error(AUSM_ERR_ARG, "NULL argument");
  if (kind < 0 || kind > 9) return set_error(AUSM_ERR_ARG, "unknown field %d", kind);
  if (!c->has_state) return set_error(AUSM_ERR_ORDER, "no state");
  hipSetDevice(c->device);
  float* dout = c->stage_buf;                       // N x M
  float* dfs = c->stage_buf + (size_t)c->M * c->N;  // one float after it
  if (fp64(c))
    CHECK_HIP(launch_derived_field(c->q64[0], c->dpar64, c->M, c->N, c->P, c->PS, kind, dout, dfs, c->stream), "derived field");
  else
    CHECK_HIP(launch_derived_field(c->q[0], c->dpar, c->M, c->N, c->P, c->PS, kind, dout, dfs, c->stream), "derived field");
  CHECK_HIP(hipMemcpyAsync(out, dout, sizeof(float) * (size_t)c->M * c->N, hipMemcpyDeviceToHost, c->stream), "derived field");
  CHECK_HIP(hipMemcpyAsync(c->pinned, dfs, sizeof(float), hipMemcpyDeviceToHost, c->stream), "derived field");
  CHECK_HIP(hipStreamSynchronize(c->stream), "derived field");
  if (freestream_value) *freestream_value = *(float*)c->pinned;
  return AUSM_OK;
}

int ausm_residual_norm(ausm_ctx* c, double* l2) {
  if (!c || !l2) return set_error(AUSM_ERR_ARG, "NULL argument");
  if (!c->has_state) return set_error(AUSM_ERR_ORDER, "no state");
  hipSetDevice(c->device);
  // a stage-1 launch into the scratch buffer Q1 accumulates the per-block
  // density-residual sums of the current solution
  const int ty = current_rows(c);
  CHECK_HIP(stage_launch(c, ty, 1, 0, 0, 1, false, false, true), "residual norm");
  CHECK_HIP(launch_reduce_sum_f64(c->res_part, c->N * stage_tiles_x(c->M), c->scal_d, c->stream), "residual norm");
  CHECK_HIP(hipMemcpyAsync(c->pinned, c->scal_d, sizeof(double), hipMemcpyDeviceToHost, c->stream), "residual norm");
  CHECK_HIP(hipStreamSynchronize(c->stream), "residual norm");
  *l2 = *(double*)c->pinned;
  return AUSM_OK;
}

int ausm_profile_stages(ausm_ctx* c, int reps, float* ms3) {
  if (!c || !ms3 || reps < 1) return set_error(AUSM_ERR_ARG, "bad profiling arguments");
  if (!c->has_state) return set_error(AUSM_ERR_ORDER, "no state");
  hipSetDevice(c->device);
  const int ty = current_rows(c);
  hipEvent_t e0, e1;
  CHECK_HIP(hipEventCreate(&e0), "event");
  CHECK_HIP(hipEventCreate(&e1), "event");
  // compensated stage 3 updates lo in place: point it at the scratch
  // accumulator while timing
  struct LoRestore {
    ausm_ctx* c;
    float* keep;
    ~LoRestore() { c->lo = keep; }
  } lo_restore{c, c->lo};
  c->lo = c->acc;
  for (int st = 1; st <= 3; ++st) {
    // scratch outputs only; the residual norm is formed on check steps only
    const int in = st == 1 ? 0 : (st == 2 ? 1 : 2);
    const int out = st == 2 ? 2 : 1;
    double total = 0.0;
    // one warm-up launch, then `reps` back-to-back launches between one
    // event pair (the per-launch event overhead is amortised)
    CHECK_HIP(stage_launch(c, ty, st, in, 0, out, false, st == 3, false), "stage");
    CHECK_HIP(hipEventRecord(e0, c->stream), "event");
    for (int k = 0; k < reps; ++k)
      CHECK_HIP(stage_launch(c, ty, st, in, 0, out, false, st == 3, false), "stage");
    CHECK_HIP(hipEventRecord(e1, c->stream), "event");
    CHECK_HIP(hipEventSynchronize(e1), "event");
    float ms = 0.0f;
    CHECK_HIP(hipEventElapsedTime(&ms, e0, e1), "event");
    total = ms;
    ms3[st - 1] = (float)(total / reps);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // the stage-3 launches above overwrote the wave-speed partials: recompute
  // them from the (unchanged) solution so that the next step is unaffected
  if (c->ws_fresh) {
    const int r = refresh_ws(c);
    if (r != AUSM_OK) return r;
  }
  CHECK_HIP(hipStreamSynchronize(c->stream), "profile");
  return AUSM_OK;
}

}  // extern "C"

int ausm_check_arithmetic(int device, long long* mismatches) {
  if (!mismatches) return set_error(AUSM_ERR_ARG, "mismatches is NULL");
  int ndev = 0;
  CHECK_HIP(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
  if (device < 0 || device >= ndev)
    return set_error(AUSM_ERR_ARG, "device %d not present (%d devices)", device, ndev);
  CHECK_HIP(hipSetDevice(device), "hipSetDevice");
  const int r = arithmetic_check(mismatches);
  if (r == AUSM_OK) device_arithmetic_store(device, *mismatches);
  return r;
}
