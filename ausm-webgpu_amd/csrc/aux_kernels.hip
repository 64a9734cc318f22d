// Setup, epilogue and layout kernels of libausm:
//   * FP64 elliptic grid generation: linear interpolation
//     (src/shaders/grid.js:15-28), Jacobi sweeps of the Winslow / TTM
//     equations with P = Q = 0 (grid.js:44-110) and a red-black SOR variant;
//   * wave speeds + per-workgroup maxima for the CFL time step
//     (src/shaders/ausm.js:617-630, 535-582);
//   * surface forces summed in FP64 (src/shaders/graphs.js:16-101);
//   * FP64 reduction of the per-workgroup residual sums;
//   * conversions between the reference's AoS texture layout and the SoA
//     planes the stage kernel streams, with ghost rows formed as the
//     boundary pass (ausm.js:34-120) forms them.
#include "aux_kernels.hpp"
#include "reduce.hpp"

namespace ausm {

// --------------------------------------------------------------------------
// grid
// --------------------------------------------------------------------------
__global__ void grid_interp_kernel(double* gx, double* gy, int M, int N, int P) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y + 1;
  if (x >= M || y >= N) return;
  const double t = (double)y / (double)N;
  const double ox = gx[x], oy = gy[x];
  const double bx = gx[N * P + x], by = gy[N * P + x];
  gx[y * P + x] = ox * (1.0 - t) + bx * t;
  gy[y * P + x] = oy * (1.0 - t) + by * t;
}

// One Winslow/TTM point update with P = Q = 0 (shaders/grid.js:79-108) on
// the point's 3 x 3 stencil X, Y (0 centre, 1 left, 2 right, 3 down, 4 up,
// 5 down-left, 6 down-right, 7 up-left, 8 up-right).  Returns false when the
// local Jacobian is not negative (point kept).
__device__ __forceinline__ bool ttm_update(const double (&X)[9], const double (&Y)[9], double& nx,
                                           double& ny) {
  const double xix = 0.5 * (X[2] - X[1]), xiy = 0.5 * (Y[2] - Y[1]);
  const double etx = 0.5 * (X[4] - X[3]), ety = 0.5 * (Y[4] - Y[3]);
  const double xex = 0.25 * (((X[8] + X[5]) - X[7]) - X[6]);
  const double xey = 0.25 * (((Y[8] + Y[5]) - Y[7]) - Y[6]);
  const double alpha = etx * etx + ety * ety;
  const double gam = xix * xix + xiy * xiy;
  const double beta = xix * etx + xiy * ety;
  const double jac = xix * ety - xiy * etx;
  const double den = 2.0 * (alpha + gam);
  const double npx =
      (alpha * ((X[1] + X[2]) + 0.0 * xix) + gam * ((X[3] + X[4]) + 0.0 * etx) - (2.0 * beta) * xex) /
      den;
  const double npy =
      (alpha * ((Y[1] + Y[2]) + 0.0 * xiy) + gam * ((Y[3] + Y[4]) + 0.0 * ety) - (2.0 * beta) * xey) /
      den;
  if (jac < 0.0) {
    // mix(center, newPos, omega = 1)
    nx = X[0] * (1.0 - 1.0) + npx * 1.0;
    ny = Y[0] * (1.0 - 1.0) + npy * 1.0;
    return true;
  }
  nx = X[0];
  ny = Y[0];
  return false;
}

// the same on the point's stencil in the vertex planes (i wraps at the seam)
__device__ __forceinline__ bool ttm_point(const double* gx, const double* gy, int M, int P, int x,
                                          int y, double& nx, double& ny) {
  const int l = (x == 0) ? M - 1 : x - 1;
  const int r = (x + 1 == M) ? 0 : x + 1;
  const int u = (y + 1) * P, c = y * P, d = (y - 1) * P;
  const double X[9] = {gx[c + x], gx[c + l], gx[c + r], gx[d + x], gx[u + x],
                       gx[d + l], gx[d + r], gx[u + l], gx[u + r]};
  const double Y[9] = {gy[c + x], gy[c + l], gy[c + r], gy[d + x], gy[u + x],
                       gy[d + l], gy[d + r], gy[u + l], gy[u + r]};
  return ttm_update(X, Y, nx, ny);
}

__global__ void grid_jacobi_kernel(const double* gx, const double* gy, double* ox, double* oy, int M,
                                   int N, int P) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= M || y > N) return;
  if (y == 0 || y == N) {
    ox[y * P + x] = gx[y * P + x];
    oy[y * P + x] = gy[y * P + x];
    return;
  }
  double nx, ny;
  ttm_point(gx, gy, M, P, x, y, nx, ny);
  ox[y * P + x] = nx;
  oy[y * P + x] = ny;
}

// kGridSweeps sweeps in one launch: Jacobi sweeps, or (SOR) red-black
// half-sweeps starting with colour color0.  A workgroup loads its tile of
// kGridTX x kGridTY vertices with a halo of kGridSweeps vertices on every
// side (columns wrap around the seam; rows stop at the body and far-field
// rows, which never move) into LDS and sweeps it there: sweep s updates the
// vertices whose whole stencil was valid after sweep s - 1, a region that
// shrinks by one vertex per sweep on every side except against a fixed
// row, so that after kGridSweeps sweeps the tile itself is exact.  Every
// vertex update sees the values a sweep-by-sweep run gives it, through the
// same arithmetic: the result is bit-for-bit that of kGridSweeps launches of
// grid_jacobi_kernel, or of grid_sor_kernel with alternating colours.
constexpr int kGridSweeps = 4, kGridTX = 32, kGridTY = 8;
constexpr int kGridLX = kGridTX + 2 * kGridSweeps, kGridLY = kGridTY + 2 * kGridSweeps;

template <bool SOR>
__global__ __launch_bounds__(256) void grid_block_kernel(const double* gx, const double* gy,
                                                         double* ox, double* oy, int M, int N, int P,
                                                         int tiles_x, int color0, double omega) {
  __shared__ double sx[2][kGridLY][kGridLX], sy[2][kGridLY][kGridLX];
  const int x0 = (blockIdx.x % tiles_x) * kGridTX;
  const int y0 = (blockIdx.x / tiles_x) * kGridTY;
  const int ga = max(y0 - kGridSweeps, 0);                // first row in LDS
  const int gb = min(y0 + kGridTY + kGridSweeps, N + 1);  // one past the last
  for (int t = threadIdx.x; t < (gb - ga) * kGridLX; t += 256) {
    const int r = t / kGridLX, c = t - r * kGridLX;
    int x = (x0 - kGridSweeps + c) % M;
    if (x < 0) x += M;
    const size_t g = (size_t)(ga + r) * (size_t)P + (size_t)x;
    sx[0][r][c] = gx[g];
    sy[0][r][c] = gy[g];
  }
  __syncthreads();
  // rows [va, vb) (grid numbering) hold valid values after each sweep
  int va = ga, vb = gb;
#pragma unroll
  for (int s = 1; s <= kGridSweeps; ++s) {
    const int cur = (s - 1) & 1, nxt = s & 1;
    const int na = ga == 0 ? 0 : va + 1;
    const int nb = gb == N + 1 ? N + 1 : vb - 1;
    const int w = kGridLX - 2 * s;  // columns [s, kGridLX - s)
    for (int t = threadIdx.x; t < (nb - na) * w; t += 256) {
      const int rr = t / w;
      const int c = s + (t - rr * w);
      const int gr = na + rr, r = gr - ga;
      bool relax = gr != 0 && gr != N;
      if constexpr (SOR) {
        // this half-sweep's colour, on the vertex's own column of the ring
        int x = (x0 - kGridSweeps + c) % M;
        if (x < 0) x += M;
        relax = relax && ((x + gr) & 1) == ((color0 + s - 1) & 1);
      }
      double nx = sx[cur][r][c], ny = sy[cur][r][c];
      if (relax) {
        const double X[9] = {sx[cur][r][c],         sx[cur][r][c - 1],     sx[cur][r][c + 1],
                             sx[cur][r - 1][c],     sx[cur][r + 1][c],     sx[cur][r - 1][c - 1],
                             sx[cur][r - 1][c + 1], sx[cur][r + 1][c - 1], sx[cur][r + 1][c + 1]};
        const double Y[9] = {sy[cur][r][c],         sy[cur][r][c - 1],     sy[cur][r][c + 1],
                             sy[cur][r - 1][c],     sy[cur][r + 1][c],     sy[cur][r - 1][c - 1],
                             sy[cur][r - 1][c + 1], sy[cur][r + 1][c - 1], sy[cur][r + 1][c + 1]};
        if constexpr (SOR) {
          double tx, ty;
          if (ttm_update(X, Y, tx, ty)) {
            nx = nx + omega * (tx - nx);
            ny = ny + omega * (ty - ny);
          }
        } else {
          ttm_update(X, Y, nx, ny);
        }
      }
      sx[nxt][r][c] = nx;
      sy[nxt][r][c] = ny;
    }
    va = na;
    vb = nb;
    __syncthreads();
  }
  constexpr int fin = kGridSweeps & 1;
  const int rows_out = min(y0 + kGridTY, N + 1) - y0;
  for (int t = threadIdx.x; t < rows_out * kGridTX; t += 256) {
    const int rr = t / kGridTX, cc = t - rr * kGridTX;
    const int x = x0 + cc;
    if (x >= M) continue;
    const int gr = y0 + rr;
    const size_t g = (size_t)gr * (size_t)P + (size_t)x;
    ox[g] = sx[fin][gr - ga][kGridSweeps + cc];
    oy[g] = sy[fin][gr - ga][kGridSweeps + cc];
  }
}

// Red-black over-relaxation of the same update.  A half-sweep relaxes the
// interior vertices of one colour ((x + y) & 1) from the values after the
// previous half-sweep (gx, gy) and copies every other vertex, writing
// ox, oy: the stencil's diagonal neighbours have the vertex's own colour, so
// relaxing in place would let a vertex read neighbours that the same
// half-sweep may or may not have updated yet.
__global__ void grid_sor_kernel(const double* gx, const double* gy, double* ox, double* oy, int M,
                                int N, int P, int color, double omega) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= M || y > N) return;
  const size_t c = (size_t)y * (size_t)P + (size_t)x;
  double vx = gx[c], vy = gy[c];
  if (y > 0 && y < N && ((x + y) & 1) == color) {
    double nx, ny;
    if (ttm_point(gx, gy, M, P, x, y, nx, ny)) {
      vx = vx + omega * (nx - vx);
      vy = vy + omega * (ny - vy);
    }
  }
  ox[c] = vx;
  oy[c] = vy;
}

__global__ void f64_to_f32_kernel(const double* a, float* b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// area and face lengths (left, right, up, down) of every cell, AoS, from the
// FP32 vertices (shaders/grid.js:128-147)
__global__ void metrics_kernel(const float* vx, const float* vy, int M, int N, int P, float* area,
                               float* flen) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= M || y >= N) return;
  const int r = (x + 1 == M) ? 0 : x + 1;
  const float ax = vx[y * P + x], ay = vy[y * P + x];
  const float bx = vx[y * P + r], by = vy[y * P + r];
  const float cx = vx[(y + 1) * P + x], cy = vy[(y + 1) * P + x];
  const float dx = vx[(y + 1) * P + r], dy = vy[(y + 1) * P + r];
  area[y * M + x] = 0.5f * fabsf((dx - ax) * (cy - by) - (cx - bx) * (dy - ay));
  float* f = flen + ((size_t)y * M + x) * 4;
  float ex, ey;
  ex = ax - cx; ey = ay - cy; f[0] = sqrtf(ex * ex + ey * ey);
  ex = bx - dx; ey = by - dy; f[1] = sqrtf(ex * ex + ey * ey);
  ex = cx - dx; ey = cy - dy; f[2] = sqrtf(ex * ex + ey * ey);
  ex = ax - bx; ey = ay - by; f[3] = sqrtf(ex * ex + ey * ey);
}

// --------------------------------------------------------------------------
// wave speeds
// --------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ R nan_max2(R m, R x) { return (x > m || x != x) ? x : m; }

template <typename R>
__global__ void wave_speed_kernel(const R* q, const R* vx, const R* vy, const DevParamsT<R>* par,
                                  int M, int N, int P, int PS, R* ws, R* part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  R w = R(0);
  if (i < M * N) {
    const int y = i / M, x = i - y * M;
    const int r = (x + 1 == M) ? 0 : x + 1;
    const R ax = vx[y * P + x], ay = vy[y * P + x];
    const R bx = vx[y * P + r], by = vy[y * P + r];
    const R cx = vx[(y + 1) * P + x], cy = vy[(y + 1) * P + x];
    const R dx = vx[(y + 1) * P + r], dy = vy[(y + 1) * P + r];
    const R area = R(0.5) * fabs((dx - ax) * (cy - by) - (cx - bx) * (dy - ay));
    R ex, ey, fl[4];
    ex = ax - cx; ey = ay - cy; fl[0] = sqrt(ex * ex + ey * ey);
    ex = bx - dx; ey = by - dy; fl[1] = sqrt(ex * ex + ey * ey);
    ex = cx - dx; ey = cy - dy; fl[2] = sqrt(ex * ex + ey * ey);
    ex = ax - bx; ey = ay - by; fl[3] = sqrt(ex * ex + ey * ey);
    const int o = y * P + x;
    const R rho = q[o];
    const R u = q[PS + o] / rho, v = q[2 * PS + o] / rho;
    const R vv = u * u + v * v;
    const R p = (q[3 * PS + o] - (R(0.5) * rho) * vv) * (par->gamma - R(1));
    const R T = p / rho;
    const R c = sqrt(par->gamma * T);
    const R speed = sqrt(vv);
    const R S = ((fl[0] + fl[1]) + fl[2]) + fl[3];
    w = ((speed + c) * S) / (R(2) * area);
    if (ws) ws[i] = w;
  }
  __shared__ R sm[256];
  sm[threadIdx.x] = w;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] = nan_max2(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

template <typename R>
__global__ void reduce_max_kernel(const R* part, int n, R* out) {
  const R m = strided_max(part, n, (int)threadIdx.x, (int)blockDim.x);
  __shared__ R sm[256];
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] = nan_max2(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sm[0];
}

// square root of the sum of FP64 partials, in a fixed order (strided
// per-thread sums, then a fixed tree): deterministic for a given n
__global__ void reduce_sum_f64_kernel(const double* part, int n, double* out) {
  __shared__ double sm[256];
  const double s = strided_sum(part, n, (int)threadIdx.x, 256);
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sm[threadIdx.x] += sm[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sqrt(sm[0]);
}

// --------------------------------------------------------------------------
// forces (graphs.js:16-101): per-segment p*l*(-n) from row-1 pressure,
// summed in FP64 in index order
// --------------------------------------------------------------------------
template <typename R>
__global__ void forces_kernel(const R* q, const R* vx, const R* vy, const DevParamsT<R>* par, int M,
                              int PS, double ref_len, double* seg, double* out, double* cp) {
  const double Vx = par->vx, Vy = par->vy;
  const double qinf = 0.5 * (double)par->inRho * (Vx * Vx + Vy * Vy);
  for (int i = threadIdx.x; i < M; i += blockDim.x) {
    const R rho = q[i];
    const R u = q[PS + i] / rho, v = q[2 * PS + i] / rho;
    const R p = (q[3 * PS + i] - (R(0.5) * rho) * (u * u + v * v)) * (par->gamma - R(1));
    const int r = (i + 1 == M) ? 0 : i + 1;
    const R ex = vx[r] - vx[i], ey = vy[r] - vy[i];
    const R fl = sqrt(ex * ex + ey * ey);
    const R nx = (vy[r] - vy[i]) / fl, ny = (vx[i] - vx[r]) / fl;
    const R pf = p * fl;
    seg[2 * i] = (double)(pf * -nx);
    seg[2 * i + 1] = (double)(pf * -ny);
    if (cp) cp[i] = ((double)p - (double)par->inP) / qinf;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double fx = 0.0, fy = 0.0;
    for (int i = 0; i < M; ++i) {
      fx += seg[2 * i];
      fy += seg[2 * i + 1];
    }
    const double vn = sqrt(Vx * Vx + Vy * Vy);
    const double ux = Vx / vn, uy = Vy / vn;
    out[0] = fx;
    out[1] = fy;
    out[2] = (fx * -uy + fy * ux) / (qinf * ref_len);
    out[3] = (fx * ux + fy * uy) / (qinf * ref_len);
  }
}

// --------------------------------------------------------------------------
// layouts
// --------------------------------------------------------------------------
// reference layout (N+3) x M x 4, rows 1..N -> SoA planes
template <typename R>
__global__ void aos_to_soa_kernel(const R* aos, R* q, int M, int N, int P, int PS) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= M || y >= N) return;
  const R* s = aos + ((size_t)(y + 1) * M + x) * 4;
  const int o = y * P + x;
  q[o] = s[0];
  q[PS + o] = s[1];
  q[2 * PS + o] = s[2];
  q[3 * PS + o] = s[3];
}

// SoA planes -> reference layout, ghost rows as the boundary pass makes them
template <typename R>
__global__ void soa_to_aos_kernel(const R* q, const R* vx, const R* vy, const DevParamsT<R>* par,
                                  int M, int N, int P, int PS, R* aos) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;  // 0 .. N+2
  if (x >= M || y > N + 2) return;
  const int r = (x + 1 == M) ? 0 : x + 1;
  ConsT<R> g;
  auto ld = [&](int j) {
    const int o = j * P + x;
    return ConsT<R>{q[o], q[PS + o], q[2 * PS + o], q[3 * PS + o]};
  };
  if (y >= 1 && y <= N) {
    g = ld(y - 1);
  } else if (y == 0) {
    R nx, ny, len;
    normal_exact(vx[x] - vx[r], vy[x] - vy[r], nx, ny, len);
    g = wall_ghost(ld(0), nx, ny);
  } else {
    R nx, ny, len;
    normal_exact(vx[N * P + x] - vx[N * P + r], vy[N * P + x] - vy[N * P + r], nx, ny, len);
    g = far_ghost(*par, ld(N - 1), nx, ny, y == N + 2);
  }
  R* d = aos + ((size_t)y * M + x) * 4;
  d[0] = g.r;
  d[1] = g.mu;
  d[2] = g.mv;
  d[3] = g.e;
}

// residual planes -> N x M x 4
template <typename R>
__global__ void planes_to_aos_kernel(const R* q, int M, int N, int P, int PS, R* aos) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= M || y >= N) return;
  const int o = y * P + x;
  R* d = aos + ((size_t)y * M + x) * 4;
  d[0] = q[o];
  d[1] = q[PS + o];
  d[2] = q[2 * PS + o];
  d[3] = q[3 * PS + o];
}

// uniform state: rest (ausm.js:14-16) or inState
template <typename R>
__global__ void fill_state_kernel(R* q, const DevParamsT<R>* par, int M, int N, int P, int PS,
                                  int freestream) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= M || y >= N) return;
  const int o = y * P + x;
  if (freestream) {
    q[o] = par->inState[0];
    q[PS + o] = par->inState[1];
    q[2 * PS + o] = par->inState[2];
    q[3 * PS + o] = par->inState[3];
  } else {
    q[o] = par->inRho;
    q[PS + o] = R(0);
    q[2 * PS + o] = R(0);
    q[3 * PS + o] = par->inP / (par->gamma - R(1));
  }
}

// compensated FP32 pair (hi + lo) <-> FP64, interior rows, 4 planes
__global__ void comp_to_f64_kernel(const float* hi, const float* lo, double* q, int M, int N, int P,
                                   int PS) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= M || y >= N) return;
  const int o = y * P + x;
  for (int k = 0; k < 4; ++k) q[k * PS + o] = (double)hi[k * PS + o] + (double)lo[k * PS + o];
}
__global__ void f64_to_comp_kernel(const double* q, float* hi, float* lo, int M, int N, int P,
                                   int PS) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= M || y >= N) return;
  const int o = y * P + x;
  for (int k = 0; k < 4; ++k) {
    const double v = q[k * PS + o];
    const float h = (float)v;
    hi[k * PS + o] = h;
    lo[k * PS + o] = (float)(v - (double)h);
  }
}

// state planes between precisions (interior rows, 4 planes)
template <typename A, typename B>
__global__ void convert_planes_kernel(const A* a, B* b, int M, int N, int P, int PS) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= M || y >= N) return;
  const int o = y * P + x;
  for (int k = 0; k < 4; ++k) b[k * PS + o] = (B)a[k * PS + o];
}

// --------------------------------------------------------------------------
// derived display fields (visualization pass, ausm.js:586-710): the scalar
// before the colour map; mode = the pass's simDisplayMode (schlieren,
// vorticity, density, temperature, pressure, Mach, |v - V_inf|, entropy,
// total pressure).  Block 0's thread 0 also writes the free-stream value.
// --------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ R dexp(R x) { if constexpr (sizeof(R) == 4) return expf(x); else return exp(x); }
template <typename R>
__device__ __forceinline__ R dlog(R x) { if constexpr (sizeof(R) == 4) return logf(x); else return log(x); }
template <typename R>
__device__ __forceinline__ R dpow(R x, R y) { if constexpr (sizeof(R) == 4) return powf(x, y); else return pow(x, y); }

template <typename R>
__global__ void derived_field_kernel(const R* q, const DevParamsT<R>* par, int M, int N, int P,
                                     int PS, int mode, float* out, float* fs_out) {
  const R gamma = par->gamma;
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
    const R inVel = sqrt(par->vx * par->vx + par->vy * par->vy);
    const R inTemp = par->inP / par->inRho;
    const R inMach = inVel / sqrt(gamma * inTemp);
    R fs;
    switch (mode) {
      case 0: fs = R(1); break;
      case 1: case 7: fs = R(0); break;
      case 3: fs = inTemp; break;
      case 4: fs = par->inP; break;
      case 5: fs = R(0.5); break;
      case 6: fs = inMach; break;
      case 8: fs = dlog(par->inP / dpow(par->inRho, gamma)); break;
      case 9: fs = par->inP * dpow(R(1) + R(0.5) * (gamma - R(1)) * inMach * inMach, gamma / (gamma - R(1))); break;
      default: fs = par->inRho;
    }
    *fs_out = (float)fs;
  }
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= M || y >= N) return;
  auto ld = [&](int j, int i) {
    const int o = j * P + i;
    return ConsT<R>{q[o], q[PS + o], q[2 * PS + o], q[3 * PS + o]};
  };
  const ConsT<R> c = ld(y, x);
  const R rho = c.r;
  const R vx = c.mu / rho, vy = c.mv / rho;
  const R vv = vx * vx + vy * vy;
  const R p = (c.e - (R(0.5) * rho) * vv) * (gamma - R(1));
  const R T = p / rho;
  const R a = sqrt(gamma * T);
  const R speed = sqrt(vv);
  R v;
  if (mode <= 1) {
    const ConsT<R> L = ld(y, x == 0 ? M - 1 : x - 1);
    const ConsT<R> Rr = ld(y, x + 1 == M ? 0 : x + 1);
    const ConsT<R> D = ld(max(y - 1, 0), x);
    const ConsT<R> U = ld(min(y + 1, N - 2), x);
    if (mode == 0) {
      const R gx = Rr.r - L.r, gy = U.r - D.r;
      v = dexp(-sqrt(gx * gx + gy * gy) * R(2));
    } else {
      v = (Rr.mv / Rr.r - L.mv / L.r) - (U.mu / U.r - D.mu / D.r);
    }
  } else {
    switch (mode) {
      case 3: v = T; break;
      case 4: v = p; break;
      case 5: case 6: v = speed / a; break;
      case 7: {
        const R dx = vx - par->vx, dy = vy - par->vy;
        v = sqrt(dx * dx + dy * dy);
        break;
      }
      case 8: v = dlog(p / dpow(rho, gamma)); break;
      case 9: {
        const R m = speed / a;
        v = p * dpow(R(1) + R(0.5) * (gamma - R(1)) * m * m, gamma / (gamma - R(1)));
        break;
      }
      default: v = rho;
    }
  }
  out[(size_t)y * M + x] = (float)v;
}

// --------------------------------------------------------------------------
// launchers// --------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------
static dim3 rows_grid(int M, int rows) { return dim3((M + 255) / 256, rows); }

hipError_t launch_grid_interp(double* gx, double* gy, int M, int N, int P, hipStream_t s) {
  if (N < 2) return hipSuccess;
  hipLaunchKernelGGL(grid_interp_kernel, rows_grid(M, N - 1), dim3(256), 0, s, gx, gy, M, N, P);
  return hipGetLastError();
}
hipError_t launch_grid_jacobi(const double* gx, const double* gy, double* ox, double* oy, int M,
                              int N, int P, hipStream_t s) {
  hipLaunchKernelGGL(grid_jacobi_kernel, rows_grid(M, N + 1), dim3(256), 0, s, gx, gy, ox, oy, M, N,
                     P);
  return hipGetLastError();
}
int grid_block_sweeps() { return kGridSweeps; }
hipError_t launch_grid_jacobi_block(const double* gx, const double* gy, double* ox, double* oy,
                                    int M, int N, int P, hipStream_t s) {
  const int tiles_x = (M + kGridTX - 1) / kGridTX;
  const int tiles_y = (N + 1 + kGridTY - 1) / kGridTY;
  hipLaunchKernelGGL(grid_block_kernel<false>, dim3(tiles_x * tiles_y), dim3(256), 0, s, gx, gy, ox,
                     oy, M, N, P, tiles_x, 0, 1.0);
  return hipGetLastError();
}
hipError_t launch_grid_sor_block(const double* gx, const double* gy, double* ox, double* oy, int M,
                                 int N, int P, int color0, double omega, hipStream_t s) {
  const int tiles_x = (M + kGridTX - 1) / kGridTX;
  const int tiles_y = (N + 1 + kGridTY - 1) / kGridTY;
  hipLaunchKernelGGL(grid_block_kernel<true>, dim3(tiles_x * tiles_y), dim3(256), 0, s, gx, gy, ox,
                     oy, M, N, P, tiles_x, color0, omega);
  return hipGetLastError();
}
hipError_t launch_grid_sor(const double* gx, const double* gy, double* ox, double* oy, int M,
                           int N, int P, int color, double omega, hipStream_t s) {
  hipLaunchKernelGGL(grid_sor_kernel, rows_grid(M, N + 1), dim3(256), 0, s, gx, gy, ox, oy, M, N, P,
                     color, omega);
  return hipGetLastError();
}
hipError_t launch_f64_to_f32(const double* a, float* b, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(f64_to_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, n);
  return hipGetLastError();
}
hipError_t launch_metrics(const float* vx, const float* vy, int M, int N, int P, float* area,
                          float* flen, hipStream_t s) {
  hipLaunchKernelGGL(metrics_kernel, rows_grid(M, N), dim3(256), 0, s, vx, vy, M, N, P, area, flen);
  return hipGetLastError();
}
int wave_speed_blocks(int M, int N) { return (M * N + 255) / 256; }
template <typename R>
hipError_t launch_wave_speed(const R* q, const R* vx, const R* vy, const DevParamsT<R>* par, int M,
                             int N, int P, int PS, R* ws, R* part, hipStream_t s) {
  hipLaunchKernelGGL(wave_speed_kernel<R>, dim3(wave_speed_blocks(M, N)), dim3(256), 0, s, q, vx, vy,
                     par, M, N, P, PS, ws, part);
  return hipGetLastError();
}
template <typename R>
hipError_t launch_reduce_max(const R* part, int n, R* out, hipStream_t s) {
  hipLaunchKernelGGL(reduce_max_kernel<R>, dim3(1), dim3(256), 0, s, part, n, out);
  return hipGetLastError();
}
hipError_t launch_reduce_sum_f64(const double* part, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(reduce_sum_f64_kernel, dim3(1), dim3(256), 0, s, part, n, out);
  return hipGetLastError();
}
template <typename R>
hipError_t launch_forces(const R* q, const R* vx, const R* vy, const DevParamsT<R>* par, int M,
                         int PS, double ref_len, double* seg, double* out, double* cp,
                         hipStream_t s) {
  hipLaunchKernelGGL(forces_kernel<R>, dim3(1), dim3(256), 0, s, q, vx, vy, par, M, PS, ref_len,
                     seg, out, cp);
  return hipGetLastError();
}
template <typename R>
hipError_t launch_aos_to_soa(const R* aos, R* q, int M, int N, int P, int PS, hipStream_t s) {
  hipLaunchKernelGGL(aos_to_soa_kernel<R>, rows_grid(M, N), dim3(256), 0, s, aos, q, M, N, P, PS);
  return hipGetLastError();
}
template <typename R>
hipError_t launch_soa_to_aos(const R* q, const R* vx, const R* vy, const DevParamsT<R>* par, int M,
                             int N, int P, int PS, R* aos, hipStream_t s) {
  hipLaunchKernelGGL(soa_to_aos_kernel<R>, rows_grid(M, N + 3), dim3(256), 0, s, q, vx, vy, par, M,
                     N, P, PS, aos);
  return hipGetLastError();
}
template <typename R>
hipError_t launch_planes_to_aos(const R* q, int M, int N, int P, int PS, R* aos, hipStream_t s) {
  hipLaunchKernelGGL(planes_to_aos_kernel<R>, rows_grid(M, N), dim3(256), 0, s, q, M, N, P, PS, aos);
  return hipGetLastError();
}
template <typename R>
hipError_t launch_fill_state(R* q, const DevParamsT<R>* par, int M, int N, int P, int PS,
                             int freestream, hipStream_t s) {
  hipLaunchKernelGGL(fill_state_kernel<R>, rows_grid(M, N), dim3(256), 0, s, q, par, M, N, P, PS,
                     freestream);
  return hipGetLastError();
}
hipError_t launch_comp_to_f64(const float* hi, const float* lo, double* q, int M, int N, int P,
                              int PS, hipStream_t s) {
  hipLaunchKernelGGL(comp_to_f64_kernel, rows_grid(M, N), dim3(256), 0, s, hi, lo, q, M, N, P, PS);
  return hipGetLastError();
}
hipError_t launch_f64_to_comp(const double* q, float* hi, float* lo, int M, int N, int P, int PS,
                              hipStream_t s) {
  hipLaunchKernelGGL(f64_to_comp_kernel, rows_grid(M, N), dim3(256), 0, s, q, hi, lo, M, N, P, PS);
  return hipGetLastError();
}
template <typename A, typename B>
hipError_t launch_convert_planes(const A* a, B* b, int M, int N, int P, int PS, hipStream_t s) {
  hipLaunchKernelGGL((convert_planes_kernel<A, B>), rows_grid(M, N), dim3(256), 0, s, a, b, M, N, P,
                     PS);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_derived_field(const R* q, const DevParamsT<R>* par, int M, int N, int P, int PS,
                                int mode, float* out, float* fs_out, hipStream_t s) {
  hipLaunchKernelGGL(derived_field_kernel<R>, rows_grid(M, N), dim3(256), 0, s, q, par, M, N, P, PS,
                     mode, out, fs_out);
  return hipGetLastError();
}
template hipError_t launch_derived_field<float>(const float*, const DevParamsT<float>*, int, int, int,
                                                int, int, float*, float*, hipStream_t);
template hipError_t launch_derived_field<double>(const double*, const DevParamsT<double>*, int, int,
                                                 int, int, int, float*, float*, hipStream_t);

#define AUSM_AUX_INSTANTIATE(R)                                                                   \
  template hipError_t launch_wave_speed<R>(const R*, const R*, const R*, const DevParamsT<R>*, int, \
                                           int, int, int, R*, R*, hipStream_t);                 \
  template hipError_t launch_reduce_max<R>(const R*, int, R*, hipStream_t);                      \
  template hipError_t launch_forces<R>(const R*, const R*, const R*, const DevParamsT<R>*, int,    \
                                       int, double, double*, double*, double*, hipStream_t);     \
  template hipError_t launch_aos_to_soa<R>(const R*, R*, int, int, int, int, hipStream_t);        \
  template hipError_t launch_soa_to_aos<R>(const R*, const R*, const R*, const DevParamsT<R>*, int, \
                                           int, int, int, R*, hipStream_t);                     \
  template hipError_t launch_planes_to_aos<R>(const R*, int, int, int, int, R*, hipStream_t);     \
  template hipError_t launch_fill_state<R>(R*, const DevParamsT<R>*, int, int, int, int, int,     \
                                           hipStream_t);
AUSM_AUX_INSTANTIATE(float)
AUSM_AUX_INSTANTIATE(double)
#undef AUSM_AUX_INSTANTIATE
template hipError_t launch_convert_planes<float, double>(const float*, double*, int, int, int, int,
                                                         hipStream_t);
template hipError_t launch_convert_planes<double, float>(const double*, float*, int, int, int, int,
                                                         hipStream_t);

// --------------------------------------------------------------------------
// self-check of the exact arithmetic's short division and square root
// --------------------------------------------------------------------------
// One thread per FP32 divisor d = 2^(e-127) * 1.m, e = 1 + blockIdx.y
// (d in [2^-126, 2^126)), m = blockIdx.x * 256 + threadIdx.x: recip_rn(d)
// against the IEEE quotient 1/d; each wavefront with a mismatch adds its
// count to *bad (nothing is written when every divisor agrees).
constexpr int kDivCheckExponents = 252;
__global__ __launch_bounds__(256) void division_check_kernel(unsigned long long* bad) {
  const unsigned m = blockIdx.x * 256u + threadIdx.x;
  const unsigned e = 1u + blockIdx.y;
  const float d = __uint_as_float((e << 23) | m);
  const float y = recip_rn(d);
  const float ref = 1.0f / d;  // the compiler's IEEE division
  const unsigned long long w = __ballot(__float_as_uint(y) != __float_as_uint(ref));
  if ((threadIdx.x & 63) == 0 && w != 0ull) atomicAdd(bad, (unsigned long long)__popcll(w));
}
// sqrt_rn(x) against the compiler's IEEE square root for every FP32 x in
// [2^-96, 2^128) (biased exponents 31 .. 254) and for x = +0
constexpr int kSqrtCheckFirstExponent = 31;
constexpr int kSqrtCheckExponents = 254 - kSqrtCheckFirstExponent + 1;
__global__ __launch_bounds__(256) void sqrt_check_kernel(unsigned long long* bad) {
  const unsigned m = blockIdx.x * 256u + threadIdx.x;
  const unsigned e = kSqrtCheckFirstExponent + blockIdx.y;
  float x = __uint_as_float((e << 23) | m);
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) x = 0.0f;
  const unsigned long long w = __ballot(__float_as_uint(sqrt_rn(x)) != __float_as_uint(sqrtf(x)));
  if ((threadIdx.x & 63) == 0 && w != 0ull) atomicAdd(bad, (unsigned long long)__popcll(w));
}
hipError_t launch_arithmetic_check(unsigned long long* bad, hipStream_t s) {
  hipLaunchKernelGGL(division_check_kernel, dim3((1u << 23) / 256u, kDivCheckExponents), dim3(256), 0,
                     s, bad);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(sqrt_check_kernel, dim3((1u << 23) / 256u, kSqrtCheckExponents), dim3(256), 0, s,
                     bad);
  return hipGetLastError();
}

}  // namespace ausm
