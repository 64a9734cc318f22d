// Allocation-free solver step (ausm_step_noalloc): the TVD-RK3 step of a
// caller-owned solution for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  The caller passes the vertices and the state in device
// memory, in the reference's layouts; every intermediate buffer the step
// needs (the SoA state planes Q^n, Q1, Q2, the vertex planes, the
// per-workgroup wave-speed maxima, the residual-norm partials, the dt and
// norm scalars and the parameter block) is carved from one file-scope
// device array of fixed size, `scratch`.
//
// Each GPU holds its own copy of `scratch` for the life of the process.  In
// host code the array's name is only a host-side stand-in: its device
// address, per device, comes from hipGetSymbolAddress on the first call on
// that device and is the only address the launches use.  Every call on one
// GPU shares that GPU's array and leaves the last call's contents in it, so
// calls that overlap on one GPU (other streams, other host threads) race.
// The step overwrites everything it reads from the array before reading it,
// so nothing has to start at zero.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "aux_kernels.hpp"
#include "ausm.h"
#include "ctx.hpp"
#include "kernels.hpp"

constexpr size_t kScratchBytes = size_t(AUSM_NOALLOC_SCRATCH_MB) << 20;

// every intermediate buffer of an allocation-free step call
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace ausm {
namespace {

constexpr int kMaxDevices = 64;

std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// the current device's address of `scratch` (looked up once per device,
// with that device current)
int scratch_address(int device, unsigned char** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return set_error(AUSM_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch[device] = (unsigned char*)p;
  }
  *out = g_scratch[device];
  return AUSM_OK;
}

// where each buffer of one call lies in `scratch` (byte offsets, each
// aligned to 256 B)
struct ScratchLayout {
  int P, PS, tiles_x, max_parts;
  size_t q[3], vx, vy, ws, res, scal_f, scal_d, par, check, total;
};

ScratchLayout layout_for(int M, int N, int ty) {
  ScratchLayout L{};
  L.P = (int)round_up((size_t)M, 64);
  L.PS = (int)round_up((size_t)N * L.P, 64);
  L.tiles_x = stage_tiles_x(M);
  L.max_parts = std::max(wave_speed_blocks(M, N), stage_blocks(M, N, ty));
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = round_up(off + bytes, 256);
    return at;
  };
  for (int k = 0; k < 3; ++k) L.q[k] = take(sizeof(float) * 4 * (size_t)L.PS);
  L.vx = take(sizeof(float) * (size_t)(N + 1) * L.P);
  L.vy = take(sizeof(float) * (size_t)(N + 1) * L.P);
  L.ws = take(sizeof(float) * (size_t)L.max_parts);
  L.res = take(sizeof(double) * (size_t)N * L.tiles_x);
  L.scal_f = take(sizeof(float) * 8);
  L.scal_d = take(sizeof(double) * 8);
  L.par = take(sizeof(DevParams));
  L.check = take(sizeof(unsigned long long));
  L.total = off;
  return L;
}

// the parameter block, written by a kernel from its argument (no host
// buffer has to outlive the call, and a captured graph holds no copy node)
__global__ void put_params_kernel(DevParams p, DevParams* dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = p;
}

__global__ void zero_count_kernel(unsigned long long* n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *n = 0ull;
}

// vertices in the reference's layout ((N+1) x M x 2) -> two planes of pitch
// P, zero in the padding columns
__global__ void vertex_planes_kernel(const float* xy, int M, int P, float* vx, float* vy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t j = blockIdx.y;
  if (i >= P) return;
  float x = 0.0f, y = 0.0f;
  if (i < M) {
    x = xy[(j * M + i) * 2];
    y = xy[(j * M + i) * 2 + 1];
  }
  vx[j * P + i] = x;
  vy[j * P + i] = y;
}

}  // namespace
}  // namespace ausm

using namespace ausm;

extern "C" int ausm_step_noalloc(int device, const ausm_step_desc* d, const float* vertices,
                                 float* state, int nsteps, void* stream, ausm_step_info* info) {
  if (!d || !vertices || !state) return set_error(AUSM_ERR_ARG, "NULL argument");
  const int M = d->M, N = d->N;
  if (M < 4 || N < 4 || M > (1 << 15) || N > (1 << 15))
    return set_error(AUSM_ERR_ARG, "grid size %d x %d out of range (4..32768)", M, N);
  if (nsteps < 0) return set_error(AUSM_ERR_ARG, "negative step count");
  if (d->scheme < 0 || d->scheme > 2 || d->limiter < 0 || d->limiter > 4)
    return set_error(AUSM_ERR_ARG, "unknown scheme %d or limiter %d", d->scheme, d->limiter);
  if (d->math != AUSM_MATH_EXACT && d->math != AUSM_MATH_FAST)
    return set_error(AUSM_ERR_ARG, "unknown math mode %d", d->math);
  if (d->dt_mode != AUSM_DT_GLOBAL && d->dt_mode != AUSM_DT_LOCAL)
    return set_error(AUSM_ERR_ARG, "unknown dt mode %d", d->dt_mode);
  if (!(d->cfl > 0) || !(d->maxdt >= 0) || d->steps_per_dt_refresh < 1)
    return set_error(AUSM_ERR_ARG, "bad CFL settings (cfl=%g maxdt=%g refresh=%d)", d->cfl,
                     d->maxdt, d->steps_per_dt_refresh);
  if (!(d->rho > 0) || !(d->p > 0) || !(d->gamma > 1))
    return set_error(AUSM_ERR_ARG, "bad free stream (rho=%g p=%g gamma=%g)", d->rho, d->p,
                     d->gamma);
  const int ty = d->rows_per_wave ? d->rows_per_wave : 2;
  if (ty != 1 && ty != 2 && ty != 4 && ty != 8)
    return set_error(AUSM_ERR_ARG, "rows per wavefront must be 0 (default 2), 1, 2, 4 or 8");
  const ScratchLayout L = layout_for(M, N, ty);
  if (L.total > kScratchBytes)
    return set_error(AUSM_ERR_ARG,
                     "grid %d x %d needs %zu bytes of scratch; the allocation-free step has %zu "
                     "(at most about %zu cells)",
                     M, N, L.total, kScratchBytes, kScratchBytes / 56);
  if (device < 0 || device >= kMaxDevices)
    return set_error(AUSM_ERR_ARG, "device %d out of range", device);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device >= ndev)
    return set_error(AUSM_ERR_ARG, "device %d not present", device);
  if (hipSetDevice(device) != hipSuccess) return set_error(AUSM_ERR_HIP, "hipSetDevice failed");
  unsigned char* base = nullptr;
  const int r = scratch_address(device, &base);
  if (r != AUSM_OK) return r;

  hipStream_t s = (hipStream_t)stream;
  float* q[3];
  for (int k = 0; k < 3; ++k) q[k] = (float*)(base + L.q[k]);
  float* vx = (float*)(base + L.vx);
  float* vy = (float*)(base + L.vy);
  float* ws_part = (float*)(base + L.ws);
  double* res_part = (double*)(base + L.res);
  float* scal_f = (float*)(base + L.scal_f);
  double* scal_d = (double*)(base + L.scal_d);
  DevParams* dpar = (DevParams*)(base + L.par);

#define NA_CHECK(expr, what)                                  \
  do {                                                        \
    hipError_t _e = (expr);                                   \
    if (_e != hipSuccess) return hip_fail(_e, what);          \
  } while (0)
  if (d->math == AUSM_MATH_EXACT) {
    // exact math rests on the device's short division and square root:
    // the first exact call on a device runs the self-check (as context and
    // batch creation do), counting into `scratch`, and every later call
    // consults its result; under stream capture the check cannot run
    long long bad = device_arithmetic_cached(device);
    if (bad < 0) {
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      NA_CHECK(hipStreamIsCapturing(s, &cs), "capture status");
      if (cs != hipStreamCaptureStatusNone)
        return set_error(AUSM_ERR_ORDER,
                         "exact math on device %d: run ausm_check_arithmetic(%d) once before "
                         "capturing allocation-free steps", device, device);
      unsigned long long* cnt = (unsigned long long*)(base + L.check);
      hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(64), 0, s, cnt);
      NA_CHECK(hipGetLastError(), "arithmetic check");
      NA_CHECK(launch_arithmetic_check(cnt, s), "arithmetic check");
      unsigned long long h = 0;
      NA_CHECK(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, s), "arithmetic check");
      NA_CHECK(hipStreamSynchronize(s), "arithmetic check");
      bad = (long long)h;
      device_arithmetic_store(device, bad);
    }
    const int v = device_arithmetic_status(device, bad);
    if (v != AUSM_OK) return v;
  }

  DevParams hp{};
  make_freestream(hp, d->rho, d->p, d->vx, d->vy, d->gamma);
  hp.maxdt = d->maxdt;
  hp.cfl = d->cfl;
  hp.Kp = d->K_p;
  hp.Ku = d->K_u;
  hipLaunchKernelGGL(put_params_kernel, dim3(1), dim3(64), 0, s, hp, dpar);
  NA_CHECK(hipGetLastError(), "parameters");
  hipLaunchKernelGGL(vertex_planes_kernel, dim3((L.P + 255) / 256, N + 1), dim3(256), 0, s,
                     vertices, M, L.P, vx, vy);
  NA_CHECK(hipGetLastError(), "vertex planes");
  NA_CHECK(launch_aos_to_soa(state, q[0], M, N, L.P, L.PS, s), "state planes");
  // the max wave speed of the input state gives the first dt (the post
  // pass of prepareState, main.js:649-652)
  NA_CHECK(launch_wave_speed(q[0], vx, vy, dpar, M, N, L.P, L.PS, (float*)nullptr, ws_part, s),
           "wave speeds");
  int n_ws = wave_speed_blocks(M, N);

  auto stage = [&](int st, int in, int out, bool fresh, bool refresh, bool norm) {
    StageArgs a{};
    a.qin = q[in];
    a.qn = q[0];
    a.qout = q[out];
    a.vx = vx;
    a.vy = vy;
    a.par = dpar;
    a.ws_in = ws_part;
    a.ws_out = ws_part;
    a.res_out = norm ? res_part : nullptr;
    a.maxws_cur = scal_f;
    a.n_ws_in = n_ws;
    a.M = M;
    a.N = N;
    a.P = L.P;
    a.PS = L.PS;
    a.stage = st;
    a.limiter = d->limiter;
    a.fresh = fresh;
    a.refresh = refresh;
    a.tiles_x = L.tiles_x;
    a.local_dt = d->dt_mode == AUSM_DT_LOCAL;
    return launch_stage(a, d->scheme, d->math, ty, s);
  };
  // the steps of ausm_step with a fresh dt reduction at the start and one
  // refresh every steps_per_dt_refresh steps
  const int K = d->steps_per_dt_refresh;
  int since = 0;
  bool fresh = true;
  for (int k = 0; k < nsteps; ++k) {
    const bool refresh = since + 1 >= K;
    bool f1 = fresh;
    if (f1 && n_ws > kWsInKernelMax) {
      NA_CHECK(launch_reduce_max(ws_part, n_ws, scal_f, s), "max wave speed");
      f1 = false;
    }
    NA_CHECK(stage(1, 0, 1, f1, false, k == nsteps - 1), "stage 1");
    NA_CHECK(stage(2, 1, 2, false, false, false), "stage 2");
    NA_CHECK(stage(3, 2, 0, false, refresh, false), "stage 3");
    fresh = refresh;
    since = refresh ? 0 : since + 1;
    if (refresh) n_ws = stage_blocks(M, N, ty);
  }
  // the new state, ghost rows included, in the caller's buffer
  NA_CHECK(launch_soa_to_aos(q[0], vx, vy, dpar, M, N, L.P, L.PS, state, s), "state download");
  if (info) {
    // the same read-back as ausm_step's; it synchronizes the stream
    NA_CHECK(launch_reduce_sum_f64(res_part, N * L.tiles_x, scal_d, s), "residual norm");
    NA_CHECK(launch_reduce_max(ws_part, n_ws, scal_f + 1, s), "max wave speed");
    double l2 = 0.0;
    float h[2] = {0.0f, 0.0f};
    NA_CHECK(hipMemcpyAsync(&l2, scal_d, sizeof(double), hipMemcpyDeviceToHost, s), "info");
    NA_CHECK(hipMemcpyAsync(h, scal_f, 2 * sizeof(float), hipMemcpyDeviceToHost, s), "info");
    NA_CHECK(hipStreamSynchronize(s), "step");
    info->res_l2 = nsteps > 0 ? l2 : -1.0;
    info->max_wave_speed = h[0];
    info->dt = std::fmin(hp.maxdt, hp.cfl / h[0]);
    info->nonfinite = !(std::isfinite(h[0]) && std::isfinite(h[1])) || !std::isfinite(l2);
    info->steps_done = nsteps;
    if (info->nonfinite)
      return set_error(AUSM_ERR_NONFINITE, "non-finite solution after %d steps", nsteps);
  } else {
    NA_CHECK(hipGetLastError(), "step");
  }
#undef NA_CHECK
  return AUSM_OK;
}
