// The fused RK-stage tile body shared by the single-grid kernel and the
// batched persistent kernel (see stage_kernels.hip for the design).
#pragma once
#include <type_traits>

#include "kernels.hpp"
#include "reduce.hpp"

namespace ausm {

__device__ __forceinline__ int xcd_remap(int bid, int nb) {
  // Workgroups are dealt round-robin to the 8 XCDs; give each XCD one
  // contiguous run of tiles so that halo re-reads stay in its own L2.
  const int q = nb >> 3, r = nb & 7;
  const int x = bid & 7, k = bid >> 3;
  return x * q + (x < r ? x : r) + k;
}

// Lane shifts across the whole wavefront with DPP (a VALU move, no LDS):
// wave_shr:1 gives lane i the value of lane i-1, wave_shl:1 that of lane
// i+1.  Edge lanes receive 0; they are halo lanes whose results are unused.
__device__ __forceinline__ float from_left(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_right(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
}
// doubles move as two 32-bit halves
__device__ __forceinline__ double from_left(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x138, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x138, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double from_right(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x130, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x130, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}
template <typename R>
__device__ __forceinline__ PrimT<R> from_left(PrimT<R> w) {
  return PrimT<R>{from_left(w.r), from_left(w.u), from_left(w.v), from_left(w.p)};
}
template <typename R>
__device__ __forceinline__ PrimT<R> from_right(PrimT<R> w) {
  return PrimT<R>{from_right(w.r), from_right(w.u), from_right(w.v), from_right(w.p)};
}

template <typename R>
__device__ __forceinline__ R nan_max(R m, R x) { return (x > m || x != x) ? x : m; }

// DPP move with a row-level control; lanes whose source is outside the
// pattern keep their own value (old = src), the identity of a max.
template <int CTRL>
__device__ __forceinline__ float dpp_keep(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_keep(double v) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane63(float v) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ double lane63(double v) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                          __builtin_amdgcn_readlane(__double2loint(v), 63));
}

// NaN-propagating max over the wavefront, in registers: row_shr:1,2,4,8
// leave each 16-lane row's max in its lane 15, row_bcast:15 and
// row_bcast:31 carry rows 0-2 up to lane 63, which is broadcast.  Max is
// exact, so the result does not depend on the combining order.
template <typename R>
__device__ __forceinline__ R wave_max(R v) {
  v = nan_max(v, dpp_keep<0x111>(v));  // row_shr:1
  v = nan_max(v, dpp_keep<0x112>(v));  // row_shr:2
  v = nan_max(v, dpp_keep<0x114>(v));  // row_shr:4
  v = nan_max(v, dpp_keep<0x118>(v));  // row_shr:8
  v = nan_max(v, dpp_keep<0x142>(v));  // row_bcast:15
  v = nan_max(v, dpp_keep<0x143>(v));  // row_bcast:31
  return lane63(v);
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Face flux times the face length: W_L = b, W_R = c, the stencil differences
// dL = W_L - W_L2, dF = W_R - W_L, dR = W_R2 - W_R, and the face's edge
// vector (dx, dy) with its unit normal (nx, ny) = rot90(d)/len.  Exact math
// scales the unit-normal flux by len afterwards, as the reference's residual
// does; fast math folds len into the flux.
template <int SCHEME, bool FAST, typename R>
__device__ __forceinline__ FluxT<R> face_flux(const DevParamsT<R>& P, int lim, PrimT<R> b,
                                              PrimT<R> c, PrimT<R> dL, PrimT<R> dF, PrimT<R> dR,
                                              R dx, R dy, R nx, R ny, R len,
                                              FreezeT<R> frz = FreezeT<R>{0, nullptr, nullptr}) {
  if constexpr (FAST) {
    return face_flux_fast<SCHEME>(P, lim, b, c, dL, dF, dR, nx, ny, len, -dy, dx, frz);
  } else {
    const FluxT<R> G = face_flux_exact<SCHEME>(P, lim, b, c, dL, dF, dR, nx, ny, frz);
    return FluxT<R>{G.f0 * len, G.f1 * len, G.f2 * len, G.f3 * len};
  }
}

template <bool FAST, typename R>
__device__ __forceinline__ void face_normal(R dx, R dy, R& nx, R& ny, R& len) {
  if constexpr (FAST) normal_fast(dx, dy, nx, ny, len);
  else normal_exact(dx, dy, nx, ny, len);
}

// CFL wave speed of one cell, (|u| + c) * (l_L + l_R + l_U + l_D) / (2 A)
// (ausm.js:617-630); S is the sum of the four face lengths.
template <bool FAST, typename R>
__device__ __forceinline__ R cell_wave_speed(const DevParamsT<R>& P, R rho, R mu, R mv, R e, R S,
                                             R area) {
  if constexpr (FAST) {
    const float ir = __builtin_amdgcn_rcpf(rho);
    const float vx = mu * ir, vy = mv * ir;
    const float vv = fmaf(vx, vx, vy * vy);
    const float p = fmaf(-0.5f * rho, vv, e) * P.gm1;
    const float c = __builtin_amdgcn_sqrtf(P.gamma * p * ir);
    const float speed = __builtin_amdgcn_sqrtf(vv);
    return (speed + c) * S * __builtin_amdgcn_rcpf(2.0f * area);
  } else {
    // the density of a physical state is inside div_free's divisor range
    const R mom[2] = {mu, mv};
    R vel[2];
    div_free(mom, rho, vel);
    const R vx = vel[0], vy = vel[1];
    const R vv = vx * vx + vy * vy;
    const R p = (e - (R(0.5) * rho) * vv) * (P.gamma - R(1));
    // p and rho of a physical state, and (|u| + c) S >= c S > 0 over twice
    // the cell area, are inside div_rn's range, and gamma p / rho inside
    // sqrt_rn's
    const R T = div_rn(p, rho);
    const R c = sqrt_rn(P.gamma * T);
    const R speed = sqrt(vv);
    return div_rn((speed + c) * S, R(2) * area);
  }
}

// One wavefront's share of a workgroup tile: the strip of columns
// [60*tx, 60*tx+60) and rows [(4*tyb + wave)*TY, ... + TY) of one grid.
// Returns the wavefront's max wave speed of the new state (stage 3 with
// refresh; 0 otherwise).  All 64 lanes must call it together.  R = double
// is the all-FP64 variant of the exact arithmetic.
//
// COMP: compensated-state stages.  The same TVD-RK3 in increment form,
//   K_s = dt L(Q_{s-1}),  Q1 = Q^n + K1,  Q2 = Q^n + (K1 + K2)/4,
//   Q^{n+1} = Q^n + (K1 + K2 + 4 K3)/6,
// with Q^n held as hi + lo (qn / qout = hi, a.lo = lo) and the last
// increment added by an error-free TwoSum, so that increments below half an
// ulp of hi accumulate in lo instead of being rounded away.  a.acc carries
// K1, then K1 + K2, from stage to stage at the cell.
// Row access: every row a wavefront touches is wave-uniform, so an element
// is a uniform row pointer (scalar registers) plus the lane's 32-bit byte
// offset of its column, the saddr + voffset form of a global access (no
// per-access 64-bit address arithmetic in vector registers).
template <typename T>
__device__ __forceinline__ T ld_row(const T* row, uint32_t boff) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(row) + boff);
}
template <typename T>
__device__ __forceinline__ void st_row(T* row, uint32_t boff, T v) {
  __builtin_nontemporal_store(v, reinterpret_cast<T*>(reinterpret_cast<char*>(row) + boff));
}
// the four planes of one state row at one lane
template <typename R>
__device__ __forceinline__ ConsT<R> ld_cons(const R* row, size_t PS, uint32_t boff) {
  return ConsT<R>{ld_row(row, boff), ld_row(row + PS, boff), ld_row(row + 2 * PS, boff),
                  ld_row(row + 3 * PS, boff)};
}
template <typename R>
__device__ __forceinline__ void st_cons(R* row, size_t PS, uint32_t boff, R a, R b, R c, R d) {
  st_row(row, boff, a);
  st_row(row + PS, boff, b);
  st_row(row + 2 * PS, boff, c);
  st_row(row + 3 * PS, boff, d);
}

// The workgroup's LDS buffer for the faces its wavefronts hand up
// (kWavesPerBlock x 5 x 64 values); a kernel declares one and passes it to
// every tile it runs.
template <typename R>
using FaceLds = R[kWavesPerBlock][5][kWave];

// dt_arg is the stage's dt, or a callable that returns it; a callable is
// evaluated once, after the workgroup barrier, so that a reduction behind it
// can overlap the tile's loads and first face.  FRZ: the limiter-freezing
// record and reconstruction (a.freeze) are compiled in; only code that
// steps frozen solves instantiates it (with the run-time limiter), so the
// other variants carry neither its code nor its registers.
template <int SCHEME, bool FAST, int TY, typename R = float, bool COMP = false, int LIM = -1,
          int STAGE = -1, bool FRZ = false, typename DtF = R>
__device__ __forceinline__ R stage_tile(const DevParamsT<R>& P, const StageArgsT<R>& a, int tx,
                                        int tyb, FaceLds<R>& s_face, DtF dt_arg) {
  static_assert(!FAST || sizeof(R) == 4, "the fast arithmetic is FP32 only");
  using Prim = PrimT<R>;
  using Cons = ConsT<R>;
  using Flux = FluxT<R>;
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const int M = a.M, N = a.N, Pp = a.P;
  const size_t PS = (size_t)a.PS;
  const int c0 = tx * kOwn;
  const int own = min(kOwn, M - c0);
  int col = (c0 - 2 + lane) % M;
  if (col < 0) col += M;
  const bool owned = lane >= 2 && lane < 2 + own;
  const uint32_t bc = (uint32_t)col * (uint32_t)sizeof(R);    // this lane's column
  auto row_ptr = [&](const R* base, int j) { return base + (size_t)j * (size_t)Pp; };
  // row block of this wavefront among the workgroup's kWavesPerBlock: the
  // bottom block has one face more than the others, so the block order is
  // rotated with the tile row and the extra work does not always fall on the
  // same SIMD of a compute unit
  static_assert((kWavesPerBlock & (kWavesPerBlock - 1)) == 0, "rotation needs a power of two");
  const int rb = (wid + tyb) & (kWavesPerBlock - 1);
  const int j0 = (tyb * kWavesPerBlock + rb) * TY;
  // LIM >= 0: the limiter fixed at compile time (no limiter switch in the
  // face code); -1: a.limiter at run time
  const int lim = LIM >= 0 ? LIM : a.limiter;
  // STAGE >= 0: the RK stage fixed at compile time; -1: a.stage at run time
  const int stage = STAGE >= 0 ? STAGE : a.stage;

  R ws_acc = R(0);

  // Waves whose rows start past the grid still load (the rows are clamped,
  // so every access is in bounds) and take part in the workgroup barriers,
  // but compute and store nothing.
  const bool active = j0 < N;
  const bool full = j0 + TY <= N;  // all TY rows inside the grid
  {
    const R* __restrict__ qin = a.qin;
    // ---- prefetch: every row the tile reads, issued before any arithmetic.
    // Stage-input rows j0-2 .. j0+TY+1 (clamped to the interior; the clamped
    // copies of rows 0 and N-1 are what the ghost rows are formed from),
    // vertex rows j0 .. j0+TY and, for stages 2 and 3, Q^n at the tile.
    constexpr int NR = TY + 4;
    Cons raw[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      // slot 0 feeds only the bottom face, which the waves above the
      // workgroup's bottom row block take from LDS (rb is wave-uniform)
      if (k == 0 && rb != 0) continue;
      const int j = min(max(j0 - 2 + k, 0), N - 1);
      raw[k] = ld_cons(row_ptr(qin, j), PS, bc);
    }
    R vxr[TY + 1], vyr[TY + 1];
#pragma unroll
    for (int k = 0; k <= TY; ++k) {
      const int j = min(j0 + k, N);
      vxr[k] = ld_row(row_ptr(a.vx, j), bc);
      vyr[k] = ld_row(row_ptr(a.vy, j), bc);
    }
    Cons qnr[TY];
    if (stage >= 2) {
#pragma unroll
      for (int k = 0; k < TY; ++k) {
        qnr[k] = ld_cons(row_ptr(a.qn, min(j0 + k, N - 1)), PS, bc);
      }
    }
    // compensated stages: the increment accumulator (stages 2, 3) and the
    // low part of Q^n (stage 3) at the tile's cells
    Cons accr[COMP ? TY : 1], lor[COMP ? TY : 1];
    if constexpr (COMP) {
      if (stage >= 2) {
#pragma unroll
        for (int k = 0; k < TY; ++k) {
          const int j = min(j0 + k, N - 1);
          accr[k] = ld_cons(row_ptr(a.acc, j), PS, bc);
          if (stage == 3) lor[k] = ld_cons(row_ptr(a.lo, j), PS, bc);
        }
      }
    }
    // ghost rows, formed once per wavefront when its window reaches them:
    // the body ghost (rows -1 and -2 of the window) from interior row 1, the
    // two far-field ghosts from interior row N (ausm.js:57-117)
    auto to_prim = [&](const Cons& q) -> Prim {
      if constexpr (FAST) return to_prim_fast(q, P.gm1);
      else return to_prim_exact(q, P.gamma);
    };
    Prim gW{}, gF1{}, gF2{};
    // (fast math forms them with the hardware reciprocal and square root
    // too).  Their source rows are in the window already: row 0 is slot
    // 2 - j0 and row N-1 slot N+1-j0; vertex rows 0 and N are vertex slots
    // 0 and N-j0 when the window reaches them, else loaded.  A ghost row
    // enters only the vertical faces of its own lane, so the right-hand
    // vertex of the segment normal comes by a lane shift: lane 63, which
    // gets none, is a halo lane whose faces and updates are not used.
    if (active && j0 < 2) {
      R x0, y0;
      if (j0 == 0) {
        x0 = vxr[0];
        y0 = vyr[0];
      } else {
        x0 = ld_row(a.vx, bc);
        y0 = ld_row(a.vy, bc);
      }
      R nx, ny, len;
      face_normal<FAST>(x0 - from_right(x0), y0 - from_right(y0), nx, ny, len);
      gW = to_prim(wall_ghost<R, FAST>(j0 == 0 ? raw[2] : raw[1], nx, ny));
    }
    if (active && j0 + TY + 2 > N) {
      const int sN = N + 1 - j0;  // window slot of row N-1, in [2, TY+2]
      Cons I = raw[2];
#pragma unroll
      for (int k = 3; k <= TY + 2; ++k)
        if (k == sN) I = raw[k];
      const int vN = N - j0;  // vertex slot of row N, in [1, TY+1]
      R xN, yN;
      if (vN <= TY) {
        xN = vxr[1];
        yN = vyr[1];
#pragma unroll
        for (int k = 2; k <= TY; ++k)
          if (k == vN) {
            xN = vxr[k];
            yN = vyr[k];
          }
      } else {
        xN = ld_row(row_ptr(a.vx, N), bc);
        yN = ld_row(row_ptr(a.vy, N), bc);
      }
      R nx, ny, len;
      face_normal<FAST>(xN - from_right(xN), yN - from_right(yN), nx, ny, len);
      const Cons g1 = far_ghost<R, FAST>(P, I, nx, ny, false);
      gF1 = to_prim(g1);
      gF2 = to_prim(far_ghost_outer(P, g1, I, nx, ny));
    }
    // primitive state of window slot k (row j0-2+k)
    auto prim_of = [&](int k) -> Prim {
      const int j = j0 - 2 + k;
      if (j < 0) return gW;
      if (j == N) return gF1;
      if (j > N) return gF2;
      return to_prim(raw[k]);
    };
    // every window slot converted once, and the differences of neighbouring
    // slots, dd[k] = w[k] - w[k-1]: the vertical face over slots k-2..k+1
    // takes W_L = w[k-1], W_R = w[k] and dd[k-1], dd[k], dd[k+1], so each
    // difference serves up to three faces
    Prim w[NR], dd[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k)
      if (k > 0 || rb == 0) w[k] = prim_of(k);
#pragma unroll
    for (int k = 1; k < NR; ++k)
      if (k > 1 || rb == 0) dd[k] = pdiff(w[k], w[k - 1]);
    // limiter freezing: a frozen face loads its 12 recorded values where it
    // is evaluated, and a recording pass stores them at the lanes that own
    // the column.  Planes 0-11: vertical faces, row jf = face between rows
    // jf-1 and jf; planes 12-23: horizontal faces, row r = left face of the
    // cell.
    static_assert(!FRZ || LIM < 0, "freezing runs with the run-time limiter");
    const int frz = FRZ ? a.freeze : 0;
    const size_t FPl = psi_plane(N, Pp);
    auto get_psi = [&](size_t plane0, int row, R* ps) {
      const R* p = a.psi + plane0 + (size_t)row * (size_t)Pp;
#pragma unroll
      for (int c = 0; c < kPsiPerFace; ++c) ps[c] = ld_row(p + c * FPl, bc);
    };
    auto put_psi = [&](size_t plane0, int row, const R* rec) {
      R* p = a.psi + plane0 + (size_t)row * (size_t)Pp;
#pragma unroll
      for (int c = 0; c < kPsiPerFace; ++c) st_row(p + c * FPl, bc, rec[c]);
    };
    // vertical face over slots k-2..k+1 (face index j0-2+k) with edge
    // (ex, ey), times its length
    auto vface = [&](int k, R ex, R ey, R& len) -> Flux {
      R nx, ny;
      face_normal<FAST>(ex, ey, nx, ny, len);
      if constexpr (FRZ) {
        R ps[kPsiPerFace];
        if (frz == 2) get_psi(0, min(j0 - 2 + k, N), ps);
        const Flux G = face_flux<SCHEME, FAST>(P, lim, w[k - 1], w[k], dd[k - 1], dd[k], dd[k + 1],
                                               ex, ey, nx, ny, len, FreezeT<R>{frz, ps, ps});
        if (frz == 1 && owned) put_psi(0, j0 - 2 + k, ps);
        return G;
      } else {
        return face_flux<SCHEME, FAST>(P, lim, w[k - 1], w[k], dd[k - 1], dd[k], dd[k + 1], ex, ey,
                                       nx, ny, len);
      }
    };

    // top face of a full tile (face index j0 + TY, window slots TY..TY+3),
    // first: it is the bottom face of the wave above, which takes it from
    // s_face after the barrier
    Flux Gtop{};
    R lTop = R(0);
    if (full) {
      const R vt = vxr[TY], wt = vyr[TY];
      Gtop = vface(TY + 2, vt - from_right(vt), wt - from_right(wt), lTop);
      s_face[rb][0][lane] = Gtop.f0;
      s_face[rb][1][lane] = Gtop.f1;
      s_face[rb][2][lane] = Gtop.f2;
      s_face[rb][3][lane] = Gtop.f3;
      s_face[rb][4][lane] = lTop;
    }
    __syncthreads();
    R dt;
    if constexpr (std::is_invocable_v<DtF>) dt = dt_arg();
    else dt = dt_arg;

    if (active) {
      R vx0 = vxr[0], vy0 = vyr[0];
      R vx0r = from_right(vx0), vy0r = from_right(vy0);

      // bottom face of row j0: vertical flux at face index j0 (ausm.js:191-205);
      // above the workgroup's bottom row block it is the top face of the block
      // below (whose tile is full)
      Flux Gd;
      R lD;
      if (rb > 0) {
        Gd = Flux{s_face[rb - 1][0][lane], s_face[rb - 1][1][lane], s_face[rb - 1][2][lane],
                  s_face[rb - 1][3][lane]};
        lD = s_face[rb - 1][4][lane];
      } else {
        Gd = vface(2, vx0 - vx0r, vy0 - vy0r, lD);
      }

  #pragma unroll
      for (int t = 0; t < TY; ++t) {
        const int r = j0 + t;
        if (r >= N) break;
        const R vx1 = vxr[t + 1], vy1 = vyr[t + 1];
        const R vx1r = from_right(vx1), vy1r = from_right(vy1);

        // horizontal face at the left of cell (i, r): stencil i-2..i+1; each
        // lane forms its own difference w(i) - w(i-1), its neighbours' come
        // by shuffles
        const Prim w0 = w[t + 2];
        const Prim wl = from_left(w0);
        const Prim dx = pdiff(w0, wl);
        R nxX, nyX, lL;
        const R ex = vx1 - vx0, ey = vy1 - vy0;
        face_normal<FAST>(ex, ey, nxX, nyX, lL);
        Flux PX;
        if constexpr (FRZ) {
          R ps[kPsiPerFace];
          if (frz == 2) get_psi(kPsiPerFace * FPl, r, ps);
          PX = face_flux<SCHEME, FAST>(P, lim, wl, w0, from_left(dx), dx, from_right(dx), ex, ey, nxX,
                                       nyX, lL, FreezeT<R>{frz, ps, ps});
          if (frz == 1 && owned) put_psi(kPsiPerFace * FPl, r, ps);
        } else {
          PX = face_flux<SCHEME, FAST>(P, lim, wl, w0, from_left(dx), dx, from_right(dx), ex, ey, nxX,
                                       nyX, lL);
        }

        // vertical face at the top of cell (i, r): stencil r-1..r+2; the last
        // row of a tile (reached only when the tile is full) has it already
        R lU;
        Flux Gu;
        if (t == TY - 1) {
          Gu = Gtop;
          lU = lTop;
        } else {
          Gu = vface(t + 3, vx1 - vx1r, vy1 - vy1r, lU);
        }

        // right face of cell (i, r) is the left face of cell (i+1, r)
        const Flux PR{from_right(PX.f0), from_right(PX.f1), from_right(PX.f2), from_right(PX.f3)};
        const R lR = from_right(lL);

        // quad area (shaders/grid.js:140)
        const R area = R(0.5) * fabs((vx1r - vx0) * (vy1 - vy0r) - (vx1 - vx0r) * (vy1r - vy0));

        // residual dQ/dt = ns / A with ns = -(F_R l_R - F_L l_L + G_U l_U - G_D l_D)
        // (ausm.js:434).  Exact math divides by A expression by expression as
        // the reference does; fast math takes 1/A once, and its stage updates
        // take the increment dt L = ns (dt / A) with fused multiply-adds.
        const R ns[4] = {-(((PR.f0 - PX.f0) + Gu.f0) - Gd.f0), -(((PR.f1 - PX.f1) + Gu.f1) - Gd.f1),
                         -(((PR.f2 - PX.f2) + Gu.f2) - Gd.f2), -(((PR.f3 - PX.f3) + Gu.f3) - Gd.f3)};
        R ia = R(1);
        R rq[4];  // exact math: the four quotients ns / A (a cell area)
        if constexpr (FAST) ia = __builtin_amdgcn_rcpf(area);
        else div_free(ns, area, rq, owned);
        auto res = [&](int k) -> R {
          if constexpr (FAST) return ns[k] * ia;
          else return rq[k];
        };

        if (owned) {
          R nq[4];
          const R cq[4] = {raw[t + 2].r, raw[t + 2].mu, raw[t + 2].mv, raw[t + 2].e};
          R dt_c = dt;
          if (a.local_dt && stage >= 1) {
            // local time stepping: the cell's own CFL step from Q^n, the same
            // in all three stages
            const Cons& q0 = stage == 1 ? raw[t + 2] : qnr[t];
            const R ws0 = cell_wave_speed<FAST>(P, q0.r, q0.mu, q0.mv, q0.e, ((lL + lR) + lU) + lD,
                                                area);
            if constexpr (FAST) dt_c = fminf(P.maxdt, P.cfl * __builtin_amdgcn_rcpf(ws0));
            else dt_c = fmin(P.maxdt, P.cfl / ws0);
          }
          R kdt = R(0);  // fast math: dt / A
          if constexpr (FAST) kdt = dt_c * ia;
          // the increment dt L(Q)
          auto inc = [&](int k) -> R {
            if constexpr (FAST) return ns[k] * kdt;
            else return dt_c * res(k);
          };
          // c + dt L(Q)
          auto add_inc = [&](int k, R c) -> R {
            if constexpr (FAST) return fmaf(ns[k], kdt, c);
            else return c + dt_c * res(k);
          };
          if (stage == 0) {
            for (int k = 0; k < 4; ++k) nq[k] = res(k);
          } else if constexpr (COMP) {
            R an[4];
            if (stage == 1) {
              // K1 = dt L(Q^n);  Q1 = hi + K1
              for (int k = 0; k < 4; ++k) {
                an[k] = inc(k);
                nq[k] = cq[k] + an[k];
              }
            } else {
              const R qh[4] = {qnr[t].r, qnr[t].mu, qnr[t].mv, qnr[t].e};
              const R ac[4] = {accr[t].r, accr[t].mu, accr[t].mv, accr[t].e};
              if (stage == 2) {
                // K1 + K2;  Q2 = hi + (K1 + K2)/4
                for (int k = 0; k < 4; ++k) {
                  an[k] = add_inc(k, ac[k]);
                  if constexpr (FAST) nq[k] = fmaf(0.25f, an[k], qh[k]);
                  else nq[k] = qh[k] + R(0.25) * an[k];
                }
              } else {
                // D = (K1 + K2 + 4 K3)/6 added to hi + lo by TwoSum
                const R lw[4] = {lor[t].r, lor[t].mu, lor[t].mv, lor[t].e};
                for (int k = 0; k < 4; ++k) {
                  const R K = inc(k);
                  R D;
                  if constexpr (FAST) D = fmaf(4.0f, K, ac[k]) * (1.0f / 6.0f);
                  else D = (ac[k] + R(4) * K) / R(6);
                  const R sm = qh[k] + D;
                  const R b = sm - qh[k];
                  const R e = (qh[k] - (sm - b)) + (D - b);
                  const R tt = lw[k] + e;
                  const R hn = sm + tt;
                  const R b2 = hn - sm;
                  an[k] = (sm - (hn - b2)) + (tt - b2);  // new lo
                  nq[k] = hn;
                }
              }
            }
            R* dst = stage == 3 ? a.lo : a.acc;
            st_cons(dst + (size_t)r * (size_t)Pp, PS, bc, an[0], an[1], an[2], an[3]);
          } else if (stage == 1) {
            // Q1 = Qn + dt L(Qn)  (ausm.js:463)
            for (int k = 0; k < 4; ++k) nq[k] = add_inc(k, cq[k]);
          } else {
            const R qn[4] = {qnr[t].r, qnr[t].mu, qnr[t].mv, qnr[t].e};
            if (stage == 2) {
              // Q2 = 3/4 Qn + 1/4 (Q1 + dt L(Q1))  (ausm.js:494)
              if constexpr (FAST) {
                for (int k = 0; k < 4; ++k) nq[k] = fmaf(0.25f, add_inc(k, cq[k]), 0.75f * qn[k]);
              } else {
                for (int k = 0; k < 4; ++k) nq[k] = R(0.75) * qn[k] + R(0.25) * add_inc(k, cq[k]);
              }
            } else {
              // Qn+1 = (Qn + 2 (Q2 + dt L(Q2))) / 3  (ausm.js:525)
              if constexpr (FAST) {
                for (int k = 0; k < 4; ++k) nq[k] = fmaf(2.0f, add_inc(k, cq[k]), qn[k]) * (1.0f / 3.0f);
              } else {
                R t3[4];
                for (int k = 0; k < 4; ++k) t3[k] = qn[k] + R(2) * add_inc(k, cq[k]);
                div_free(t3, R(3), nq);
              }
            }
          }
          st_cons(a.qout + (size_t)r * (size_t)Pp, PS, bc, nq[0], nq[1], nq[2], nq[3]);
          if (stage == 3 && a.refresh) {
            const R ws = cell_wave_speed<FAST>(P, nq[0], nq[1], nq[2], nq[3],
                                               ((lL + lR) + lU) + lD, area);
            ws_acc = nan_max(ws_acc, ws);
          }
        }

        if (stage == 1 && a.res_out) {
          const double rr = owned ? (double)res(0) * (double)res(0) : 0.0;
          const double rsum = wave_sum(rr);
          if (lane == 0) a.res_out[r * a.tiles_x + tx] = rsum;
        }

        // slide the vertex window up one row
        vx0 = vx1;
        vy0 = vy1;
        vx0r = vx1r;
        vy0r = vy1r;
        Gd = Gu;
        lD = lU;
      }
    }
  }
  // the LDS faces may be rewritten by the next tile of a persistent loop
  __syncthreads();
  return ws_acc;
}

// dt = min(maxdt, cfl / max wave speed) (ausm.js:457).  On a fresh stage 1
// every wavefront reduces the per-workgroup maxima of the last refresh: the
// first kDtAhead per lane are loaded by stage_dt_load at kernel entry, in
// one batch that overlaps the tile's own loads, and reduced by
// stage_dt_finish where dt is first needed (the same order of maxima as a
// plain lane-strided loop, so dt does not depend on the split).  The host
// keeps this path to at most kWsInKernelMax maxima.
template <typename R>
struct DtAhead {
  R v[kDtAhead];
};
template <typename R>
__device__ __forceinline__ DtAhead<R> stage_dt_load(const StageArgsT<R>& a, int stage) {
  DtAhead<R> d;
  const int lane = threadIdx.x & 63;
  if (stage == 1 && a.fresh) {
#pragma unroll
    for (int k = 0; k < kDtAhead; ++k) {
      const int i = lane + k * kWave;
      d.v[k] = i < a.n_ws_in ? a.ws_in[i] : R(0);
    }
  }
  return d;
}
template <typename R>
__device__ __forceinline__ R stage_dt_finish(const DevParamsT<R>& P, const StageArgsT<R>& a,
                                             bool writer, const DtAhead<R>& d, int stage) {
  const int lane = threadIdx.x & 63;
  R maxws;
  if (stage == 1 && a.fresh) {
    R m = R(0);
#pragma unroll
    for (int k = 0; k < kDtAhead; ++k) m = nan_max(m, d.v[k]);
    m = nan_max(m, strided_max(a.ws_in, a.n_ws_in, lane + kDtAhead * kWave, kWave));
    maxws = wave_max(m);
    if (writer && threadIdx.x == 0) *a.maxws_cur = maxws;
  } else {
    maxws = *a.maxws_cur;
  }
  return fmin(P.maxdt, P.cfl / maxws);
}

// workgroup max of the wavefronts' wave speeds -> one partial per workgroup tile
template <typename R>
__device__ __forceinline__ void block_ws_partial(R ws_acc, R* out) {
  __shared__ R s_ws[kWavesPerBlock];
  const R wm = wave_max(ws_acc);
  if ((threadIdx.x & 63) == 0) s_ws[threadIdx.x >> 6] = wm;
  __syncthreads();
  if (threadIdx.x == 0) {
    R m = s_ws[0];
    for (int w = 1; w < kWavesPerBlock; ++w) m = nan_max(m, s_ws[w]);
    *out = m;
  }
  __syncthreads();
}

}  // namespace ausm
