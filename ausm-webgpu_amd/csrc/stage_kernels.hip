// Fused TVD-RK3 stage kernel.
//
// One launch replaces the reference's five passes of a stage (boundary,
// vertical flux, horizontal flux, residual, stage update:
// src/main.js:744-776, src/shaders/ausm.js:23-528):
//   * each wavefront owns a strip of 60 cells around i (64 lanes, two halo
//     lanes each side) and TY cells up j, and walks it row by row;
//   * a lane keeps the primitive states of rows r-1..r+2 of its column in
//     registers, so each cell is loaded and converted once per stage;
//   * i-neighbours (MUSCL stencil, the right face's flux, the i+1 vertex)
//     come from the adjacent lanes by wavefront shuffles;
//   * ghost rows are never stored: the body and far-field ghosts are formed
//     on the fly from interior rows 1 and N when the window reaches them;
//   * each face flux is evaluated once and multiplied by its length once;
//     the residual and the RK update follow in registers;
//   * stage 3 also forms the CFL wave speed of the new state (ausm.js:617-630)
//     and a per-workgroup max (no atomics); the next stage 1 reduces those
//     maxima itself, so dt needs no extra launch and is bitwise deterministic;
//   * stage 1 writes per-workgroup FP64 sums of the squared density residual
//     for the convergence norm.
// Memory per cell-step: state 8 x 16 B streams + one vertex per stage.
#include "kernels.hpp"

namespace ausm {

__device__ __forceinline__ int xcd_remap(int bid, int nb) {
  // Workgroups are dealt round-robin to the 8 XCDs; give each XCD one
  // contiguous run of tiles so that halo re-reads stay in its own L2.
  const int q = nb >> 3, r = nb & 7;
  const int x = bid & 7, k = bid >> 3;
  return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ float from_right(float v) { return __shfl_down(v, 1); }
__device__ __forceinline__ float from_left(float v) { return __shfl_up(v, 1); }
__device__ __forceinline__ Prim from_left(Prim w) {
  return Prim{from_left(w.r), from_left(w.u), from_left(w.v), from_left(w.p)};
}
__device__ __forceinline__ Prim from_right(Prim w) {
  return Prim{from_right(w.r), from_right(w.u), from_right(w.v), from_right(w.p)};
}

__device__ __forceinline__ float nan_max(float m, float x) { return (x > m || x != x) ? x : m; }

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = nan_max(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

template <int SCHEME, bool FAST>
__device__ __forceinline__ Flux face_flux(const DevParams& P, int lim, Prim a, Prim b, Prim c,
                                          Prim d, float nx, float ny) {
  if constexpr (FAST) return face_flux_fast<SCHEME>(P, lim, a, b, c, d, nx, ny);
  else return face_flux_exact<SCHEME>(P, lim, a, b, c, d, nx, ny);
}

template <bool FAST>
__device__ __forceinline__ void face_normal(float dx, float dy, float& nx, float& ny, float& len) {
  if constexpr (FAST) normal_fast(dx, dy, nx, ny, len);
  else normal_exact(dx, dy, nx, ny, len);
}

template <int SCHEME, bool FAST, int TY>
__global__ __launch_bounds__(256) void stage_kernel(const StageArgs a) {
  const DevParams P = *a.par;
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = tile % a.tiles_x;
  const int tyb = tile / a.tiles_x;
  const int M = a.M, N = a.N, Pp = a.P, PS = a.PS;
  const int c0 = tx * kOwn;
  const int own = min(kOwn, M - c0);
  int col = (c0 - 2 + lane) % M;
  if (col < 0) col += M;
  const int colr = (col + 1 == M) ? 0 : col + 1;
  const bool owned = lane >= 2 && lane < 2 + own;
  const int j0 = (tyb * kWavesPerBlock + wid) * TY;
  const int lim = a.limiter;
  const int stage = a.stage;

  // ---- time step: dt = min(maxdt, cfl / max wave speed) (ausm.js:457)
  float maxws;
  if (stage == 1 && a.fresh) {
    float m = 0.0f;
    for (int i = lane; i < a.n_ws_in; i += kWave) m = nan_max(m, a.ws_in[i]);
    maxws = wave_max(m);
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.maxws_cur = maxws;
  } else {
    maxws = *a.maxws_cur;
  }
  const float dt = fminf(P.maxdt, P.cfl / maxws);

  float ws_acc = 0.0f;
  double rs_acc = 0.0;

  if (j0 < N) {
    const float* __restrict__ qin = a.qin;
    auto load_cons = [&](int j) -> Cons {
      const int o = j * Pp + col;
      return Cons{qin[o], qin[PS + o], qin[2 * PS + o], qin[3 * PS + o]};
    };
    // primitive state of row j (-2 <= j <= N+1) of this lane's column;
    // interior rows also return the conserved state
    auto load_W = [&](int j, Cons& qc) -> Prim {
      Cons q;
      if (j >= 0 && j < N) {
        q = load_cons(j);
        qc = q;
      } else if (j < 0) {
        // body ghost row 0 from interior row 1 (ausm.js:57-64)
        const Cons I = load_cons(0);
        float nx, ny, len;
        normal_exact(a.vx[col] - a.vx[colr], a.vy[col] - a.vy[colr], nx, ny, len);
        q = wall_ghost(I, nx, ny);
      } else {
        // far-field ghost rows N+1, N+2 from interior row N (ausm.js:65-117)
        const Cons I = load_cons(N - 1);
        const int o = N * Pp;
        float nx, ny, len;
        normal_exact(a.vx[o + col] - a.vx[o + colr], a.vy[o + col] - a.vy[o + colr], nx, ny, len);
        q = far_ghost(P, I, nx, ny, j == N + 1);
      }
      if constexpr (FAST) return to_prim_fast(q, P.gm1);
      else return to_prim_exact(q, P.gamma);
    };

    Cons dummy, c_0, c_p1, c_p2;
    const Prim w_m2 = load_W(j0 - 2, dummy);
    Prim w_m1 = load_W(j0 - 1, dummy);
    Prim w_0 = load_W(j0, c_0);
    Prim w_p1 = load_W(j0 + 1, c_p1);
    float vx0 = a.vx[j0 * Pp + col], vy0 = a.vy[j0 * Pp + col];
    float vx0r = from_right(vx0), vy0r = from_right(vy0);

    // bottom face of row j0: vertical flux at face index j0 (ausm.js:191-205)
    Flux Gd;
    float lD;
    {
      float nx, ny;
      face_normal<FAST>(vx0 - vx0r, vy0 - vy0r, nx, ny, lD);
      const Flux G = face_flux<SCHEME, FAST>(P, lim, w_m2, w_m1, w_0, w_p1, nx, ny);
      Gd = Flux{G.f0 * lD, G.f1 * lD, G.f2 * lD, G.f3 * lD};
    }

#pragma unroll
    for (int t = 0; t < TY; ++t) {
      const int r = j0 + t;
      if (r >= N) break;
      const Prim w_p2 = load_W(r + 2, c_p2);
      const float vx1 = a.vx[(r + 1) * Pp + col], vy1 = a.vy[(r + 1) * Pp + col];
      const float vx1r = from_right(vx1), vy1r = from_right(vy1);

      // horizontal face at the left of cell (i, r): stencil i-2..i+1
      const Prim wl = from_left(w_0);
      const Prim wl2 = from_left(wl);
      const Prim wr = from_right(w_0);
      float nxX, nyX, lL;
      face_normal<FAST>(vx1 - vx0, vy1 - vy0, nxX, nyX, lL);
      const Flux FX = face_flux<SCHEME, FAST>(P, lim, wl2, wl, w_0, wr, nxX, nyX);
      const Flux PX{FX.f0 * lL, FX.f1 * lL, FX.f2 * lL, FX.f3 * lL};

      // vertical face at the top of cell (i, r): stencil r-1..r+2
      float nxY, nyY, lU;
      face_normal<FAST>(vx1 - vx1r, vy1 - vy1r, nxY, nyY, lU);
      const Flux FY = face_flux<SCHEME, FAST>(P, lim, w_m1, w_0, w_p1, w_p2, nxY, nyY);
      const Flux Gu{FY.f0 * lU, FY.f1 * lU, FY.f2 * lU, FY.f3 * lU};

      // right face of cell (i, r) is the left face of cell (i+1, r)
      const Flux PR{from_right(PX.f0), from_right(PX.f1), from_right(PX.f2), from_right(PX.f3)};
      const float lR = from_right(lL);

      // quad area (shaders/grid.js:140)
      const float area =
          0.5f * fabsf((vx1r - vx0) * (vy1 - vy0r) - (vx1 - vx0r) * (vy1r - vy0));

      // residual dQ/dt = -(F_R l_R - F_L l_L + G_U l_U - G_D l_D) / A  (ausm.js:434)
      float res[4];
      {
        const float s0 = ((PR.f0 - PX.f0) + Gu.f0) - Gd.f0;
        const float s1 = ((PR.f1 - PX.f1) + Gu.f1) - Gd.f1;
        const float s2 = ((PR.f2 - PX.f2) + Gu.f2) - Gd.f2;
        const float s3 = ((PR.f3 - PX.f3) + Gu.f3) - Gd.f3;
        if constexpr (FAST) {
          const float ia = __builtin_amdgcn_rcpf(area);
          res[0] = -s0 * ia;
          res[1] = -s1 * ia;
          res[2] = -s2 * ia;
          res[3] = -s3 * ia;
        } else {
          res[0] = (-s0) / area;
          res[1] = (-s1) / area;
          res[2] = (-s2) / area;
          res[3] = (-s3) / area;
        }
      }

      if (owned) {
        const int o = r * Pp + col;
        float nq[4];
        const float cq[4] = {c_0.r, c_0.mu, c_0.mv, c_0.e};
        if (stage == 0) {
          for (int k = 0; k < 4; ++k) nq[k] = res[k];
        } else if (stage == 1) {
          // Q1 = Qn + dt L(Qn)  (ausm.js:463)
          for (int k = 0; k < 4; ++k) nq[k] = cq[k] + dt * res[k];
          rs_acc += (double)res[0] * (double)res[0];
        } else {
          const float qn[4] = {a.qn[o], a.qn[PS + o], a.qn[2 * PS + o], a.qn[3 * PS + o]};
          if (stage == 2) {
            // Q2 = 3/4 Qn + 1/4 (Q1 + dt L(Q1))  (ausm.js:494)
            for (int k = 0; k < 4; ++k) nq[k] = 0.75f * qn[k] + 0.25f * (cq[k] + dt * res[k]);
          } else {
            // Qn+1 = (Qn + 2 (Q2 + dt L(Q2))) / 3  (ausm.js:525)
            for (int k = 0; k < 4; ++k) nq[k] = (qn[k] + 2.0f * (cq[k] + dt * res[k])) / 3.0f;
          }
        }
        a.qout[o] = nq[0];
        a.qout[PS + o] = nq[1];
        a.qout[2 * PS + o] = nq[2];
        a.qout[3 * PS + o] = nq[3];
        if (stage == 3 && a.refresh) {
          // wave speed (|u| + a) * (l_L + l_R + l_U + l_D) / (2 A)  (ausm.js:617-630)
          const float rho = nq[0];
          const float vx = nq[1] / rho, vy = nq[2] / rho;
          const float vv = vx * vx + vy * vy;
          const float p = (nq[3] - (0.5f * rho) * vv) * (P.gamma - 1.0f);
          const float T = p / rho;
          const float c = sqrtf(P.gamma * T);
          const float speed = sqrtf(vv);
          const float S = ((lL + lR) + lU) + lD;
          const float ws = ((speed + c) * S) / (2.0f * area);
          ws_acc = nan_max(ws_acc, ws);
        }
      }

      // slide the window up one row
      w_m1 = w_0;
      w_0 = w_p1;
      w_p1 = w_p2;
      c_0 = c_p1;
      c_p1 = c_p2;
      vx0 = vx1;
      vy0 = vy1;
      vx0r = vx1r;
      vy0r = vy1r;
      Gd = Gu;
      lD = lU;
    }
  }

  if ((stage == 3 && a.refresh) || (stage == 1 && a.res_out)) {
    __shared__ float s_ws[kWavesPerBlock];
    __shared__ double s_rs[kWavesPerBlock];
    const float wm = wave_max(ws_acc);
    const double rsum = wave_sum(rs_acc);
    if (lane == 0) {
      s_ws[wid] = wm;
      s_rs[wid] = rsum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      if (stage == 3) {
        float m = s_ws[0];
        for (int w = 1; w < kWavesPerBlock; ++w) m = nan_max(m, s_ws[w]);
        a.ws_out[blockIdx.x] = m;
      } else {
        double s = s_rs[0];
        for (int w = 1; w < kWavesPerBlock; ++w) s += s_rs[w];
        a.res_out[blockIdx.x] = s;
      }
    }
  }
}

int stage_tiles_x(int M) { return (M + kOwn - 1) / kOwn; }

int stage_blocks(int M, int N, int ty) {
  const int rows_per_block = ty * kWavesPerBlock;
  return stage_tiles_x(M) * ((N + rows_per_block - 1) / rows_per_block);
}

template <int SCHEME, bool FAST>
static hipError_t launch_ty(const StageArgs& a, int ty, hipStream_t s) {
  const int nb = stage_blocks(a.M, a.N, ty);
  switch (ty) {
    case 1: hipLaunchKernelGGL((stage_kernel<SCHEME, FAST, 1>), dim3(nb), dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((stage_kernel<SCHEME, FAST, 2>), dim3(nb), dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL((stage_kernel<SCHEME, FAST, 4>), dim3(nb), dim3(256), 0, s, a); break;
    case 8: hipLaunchKernelGGL((stage_kernel<SCHEME, FAST, 8>), dim3(nb), dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_stage(const StageArgs& a, int scheme, int fast, int ty, hipStream_t s) {
  if (fast) {
    switch (scheme) {
      case SCHEME_SLAU2: return launch_ty<SCHEME_SLAU2, true>(a, ty, s);
      case SCHEME_SLAU: return launch_ty<SCHEME_SLAU, true>(a, ty, s);
      case SCHEME_AUSMUP: return launch_ty<SCHEME_AUSMUP, true>(a, ty, s);
    }
  } else {
    switch (scheme) {
      case SCHEME_SLAU2: return launch_ty<SCHEME_SLAU2, false>(a, ty, s);
      case SCHEME_SLAU: return launch_ty<SCHEME_SLAU, false>(a, ty, s);
      case SCHEME_AUSMUP: return launch_ty<SCHEME_AUSMUP, false>(a, ty, s);
    }
  }
  return hipErrorInvalidValue;
}

}  // namespace ausm
