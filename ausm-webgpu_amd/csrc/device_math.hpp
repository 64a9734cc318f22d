// Device-side arithmetic of the solver step: primitive conversion, MUSCL
// reconstruction, the three AUSM-family face fluxes and the boundary ghost
// states.
//
// Two variants of the face flux:
//  * Exact — follows the reference WGSL (src/shaders/ausm.js) expression by
//    expression: IEEE-rounded division and sqrt, no FMA contraction (the
//    translation unit is built with -ffp-contract=off), the same operand
//    order.  WGSL built-ins are given fixed definitions: dot sums left to
//    right, normalize(v) = v/length(v), mix(a,b,t) = a*(1-t)+b*t,
//    reflect(e,n) = e-(2*dot(n,e))*n, step(e,x) = x>=e, sign(0) = 0.
//  * Fast — the same scheme rearranged for CDNA4: shared reciprocals
//    (v_rcp_f32), hardware sqrt, one reciprocal per limiter ratio and
//    explicit fused multiply-adds.  It differs from Exact by a few ulp per
//    operation.  Fusion is written out (fmaf), never left to the compiler's
//    contraction, so every kernel that inlines these functions (any tile
//    shape, single grid or batch) produces the same bits.
#pragma once
#include <hip/hip_runtime.h>

namespace ausm {

enum { SCHEME_SLAU2 = 0, SCHEME_SLAU = 1, SCHEME_AUSMUP = 2 };
enum { LIM_VANALBADA = 0, LIM_MINMOD = 1, LIM_VANLEER = 2, LIM_NONE = 3, LIM_VANALBADA_CLAMPED = 4 };

// Solver parameters resident in device memory so that captured graphs pick
// up changes without being rebuilt (the reference's Uniforms buffer,
// src/util.js:2-27).
template <typename R>
struct DevParamsT {
  R gamma, inRho, inP, vx, vy;
  R inState[4];
  R maxdt, cfl, Kp, Ku;
  // derived in double on the host, used by the fast path only
  R gm1;        // gamma - 1
  R ggm1;       // gamma / (gamma - 1)
  R minf_coef;  // |V|^2 / (gamma * p_inf)
  R astar_coef; // 2 (gamma - 1) / (gamma + 1)
};
using DevParams = DevParamsT<float>;

template <typename R>
struct PrimT {
  R r, u, v, p;
};
template <typename R>
struct ConsT {
  R r, mu, mv, e;
};
template <typename R>
struct FluxT {
  R f0, f1, f2, f3;
};
using Prim = PrimT<float>;
using Cons = ConsT<float>;
using Flux = FluxT<float>;

// ---------------------------------------------------------------------------
// IEEE FP32 division in fewer instructions, for operands in a known range.
// The compiler's correctly rounded division is v_div_scale (x2), v_rcp_f32,
// a Newton step on the reciprocal, a quotient with two FMA-residual
// corrections, v_div_fmas and v_div_fixup: 11 VALU instructions, of which
// the two scalings and the fixup only act on operands near the ends of the
// exponent range (denormal or huge divisor, numerator below 2^-100, quotient
// near overflow or underflow) and on zeros, infinities and NaNs.  div_rn is
// 6 instructions:
//   y = rcp(d);  y += y (1 - d y);  q = n y;  q += y (n - d q)
// The refined reciprocal y = recip_rn(d) is the correctly rounded 1/d for
// every normal d < 2^126 on gfx950: v_rcp_f32 followed by that Newton step
// was checked for every FP32 divisor in the range (ausm_check_arithmetic runs
// the check on the device; contexts and batches refuse a device that fails
// it).  With y = RN(1/d), a quotient q = RN(n y) within an ulp and an exact
// residual n - d q, one correction rounds correctly (Markstein's theorem).
// The residual is exact, and the result the IEEE quotient bit for bit,
// whenever d > 0 is normal with d < 2^126, and n = +0 or 2^-100 <= |n| with
// 2^-125 <= |n / d| < 2^96.  div_by takes the refined reciprocal, so that quotients
// with one divisor share it.  Every call site states why its operands are
// in range; elsewhere the exact path keeps the compiler's division
// (quotients whose numerators can be arbitrarily small, such as momenta and
// velocities at the front of a disturbance spreading into a uniform flow).
// A zero, infinite or NaN divisor or an infinite numerator gives NaN where
// IEEE may give an infinity: a state that reaches them is non-finite either
// way.  R = double (the FP64 solve) always divides with the compiler's
// sequence.
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ R recip_rn(R d) {
  if constexpr (sizeof(R) == 4) {
    const float y = __builtin_amdgcn_rcpf(d);
    return fmaf(fmaf(-d, y, 1.0f), y, y);
  } else {
    return R(1) / d;
  }
}
template <typename R>
__device__ __forceinline__ R div_by(R n, R d, R y) {
  if constexpr (sizeof(R) == 4) {
    const float q = n * y;
    return fmaf(fmaf(-d, q, n), y, q);
  } else {
    (void)y;
    return n / d;
  }
}
template <typename R>
__device__ __forceinline__ R div_rn(R n, R d) {
  return div_by(n, d, recip_rn(d));
}
// div_by with the residual formed as d q - n and subtracted in the
// correction.  For n != 0 the residual is exact either way, so -(d q - n)
// is n - d q and the quotient is div_by's bit for bit; for n = +-0 (q =
// +-0) d q - n is +0 and q - 0 y keeps q's sign, so a zero numerator gives
// IEEE's +-0 / d = +-0 where div_by's n - d q = +0 would turn -0 into +0.
// Same instruction count as div_by (the negations are operand modifiers).
template <typename R>
__device__ __forceinline__ R div_by_zs(R n, R d, R y) {
  if constexpr (sizeof(R) == 4) {
    const float q = n * y;
    return fmaf(-fmaf(d, q, -n), y, q);
  } else {
    (void)y;
    return n / d;
  }
}

// Quotients n[k] / d of numerators that may be arbitrarily small (momenta,
// velocities and residuals, which fall through the denormal range at the
// front of a disturbance spreading into a uniform flow) by one divisor d > 0
// with 2^-126 <= d < 2^26, the IEEE quotients bit for bit: while no lane
// holds a nonzero numerator below 2^-96 in magnitude (frexp exponent above
// -96), every quotient is div_by on one refined reciprocal (in div_rn's
// range: |n / d| > 2^-122) in the form that keeps the sign of a zero
// numerator, as IEEE keeps it (+-0 / d = +-0); otherwise the wavefront takes the compiler's
// division.  The test is wave-uniform, so the two paths never diverge
// within a wavefront; lanes with live = false (halo lanes whose quotients
// are not used) do not take part in it.  R = double: the compiler's
// division.
template <int K, typename R>
__device__ __forceinline__ void div_free(const R (&n)[K], R d, R (&q)[K], bool live = true) {
  if constexpr (sizeof(R) == 4) {
    int e = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) e = min(e, __builtin_amdgcn_frexp_expf(n[k]));
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(live && e <= -96) == 0, 1)) {
      const float y = recip_rn(d);
#pragma unroll
      for (int k = 0; k < K; ++k) q[k] = div_by_zs(n[k], d, y);
    } else {
#pragma unroll
      for (int k = 0; k < K; ++k) q[k] = n[k] / d;
    }
  } else {
#pragma unroll
    for (int k = 0; k < K; ++k) q[k] = n[k] / d;
  }
}

// IEEE FP32 square root in fewer instructions, for x = +0 or
// 2^-96 <= x < 2^128.  The compiler's correctly rounded square root is
// v_sqrt_f32 (within one ulp, not correctly rounded: about 15 % of inputs
// differ) followed by a choice among the result and its two neighbours by
// the signs of their FMA residuals x - s' s, wrapped in a scaling of inputs
// below 2^-96 and a select that passes zeros, infinities and NaN through:
// 15 instructions.  sqrt_rn is the same choice without the wrapping, 9;
// on its range the two are the same operations on the same values, and
// ausm_check_arithmetic compares them on the device for every FP32 input of
// the range.  Call sites state why their argument is in range; square roots
// of speeds (which vanish in still air) keep the compiler's.  R = double:
// the compiler's.
template <typename R>
__device__ __forceinline__ R sqrt_rn(R x) {
  if constexpr (sizeof(R) == 4) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float dn = __int_as_float(__float_as_int(s) - 1);
    const float up = __int_as_float(__float_as_int(s) + 1);
    const float rdn = fmaf(-dn, s, x);
    const float rup = fmaf(-up, s, x);
    const float r = rdn <= 0.0f ? dn : s;
    return rup > 0.0f ? up : r;
  } else {
    return sqrt(x);
  }
}

template <typename R>
__device__ __forceinline__ R wstep(R edge, R x) { return x >= edge ? R(1) : R(0); }
template <typename R>
__device__ __forceinline__ R wsign(R x) { return x > R(0) ? R(1) : (x < R(0) ? R(-1) : R(0)); }
template <typename R>
__device__ __forceinline__ R wsat(R x) { return fmin(fmax(x, R(0)), R(1)); }
template <typename R>
__device__ __forceinline__ R wmix(R a, R b, R t) { return a * (R(1) - t) + b * t; }
// mix(a, b, t) with t a step (0 or 1) at the call sites of the split
// functions below: in FP32 the select t ? b : a, which equals
// a (1 - t) + b t bit for bit whenever a and b are finite and the operand
// kept is not a -0 beside a +0 product of the other -- at these sites the
// subsonic polynomial is never -0 (its factors are bounded away from zero
// on |M| < 1) and the supersonic operand is never -0.  The equality is
// checked by brute force in FP32 on the operands these sites form
// (tests/test_exact_select.py).  FP64 keeps the product form.
template <typename R>
__device__ __forceinline__ R wsel(R a, R b, R t) {
  if constexpr (sizeof(R) == 4) return t != R(0) ? b : a;
  else return wmix(a, b, t);
}

// toPrimitive (ausm.js:145-150)
template <typename R>
__device__ __forceinline__ PrimT<R> to_prim_exact(ConsT<R> q, R gamma) {
  PrimT<R> w;
  w.r = q.r;
  // the density of a physical state is inside div_free's divisor range
  const R mom[2] = {q.mu, q.mv};
  R vel[2];
  div_free(mom, q.r, vel);
  w.u = vel[0];
  w.v = vel[1];
  const R d = w.u * w.u + w.v * w.v;
  const R p = (q.e - (R(0.5) * q.r) * d) * (gamma - R(1.0));
  w.p = fmax(R(1e-7), p);
  return w;
}

__device__ __forceinline__ Prim to_prim_fast(Cons q, float gm1) {
  Prim w;
  const float inv = __builtin_amdgcn_rcpf(q.r);
  w.r = q.r;
  w.u = q.mu * inv;
  w.v = q.mv * inv;
  const float ke = fmaf(q.mu, w.u, q.mv * w.v);
  const float p = fmaf(-0.5f, ke, q.e) * gm1;
  w.p = fmaxf(1e-7f, p);
  return w;
}

// face unit normal: normalize(d) rotated by +90 degrees; also returns |d|
// (the face length the residual uses, shaders/grid.js:143-146)
template <typename R>
__device__ __forceinline__ void normal_exact(R dx, R dy, R& nx, R& ny, R& len) {
  // an edge of a grid cell, at least 2^-48 long (dx^2 + dy^2 >= 2^-96)
  // unless the cell is degenerate
  len = sqrt_rn(dx * dx + dy * dy);
  const R d[2] = {dx, dy};
  R t[2];
  div_free(d, len, t);  // a grid edge's length
  const R tx = t[0], ty = t[1];
  nx = -ty;
  ny = tx;
}
__device__ __forceinline__ void normal_fast(float dx, float dy, float& nx, float& ny, float& len) {
  const float dd = fmaf(dx, dx, dy * dy);
  const float rs = __builtin_amdgcn_rsqf(dd);
  len = dd * rs;
  nx = -dy * rs;
  ny = dx * rs;
}

// ---------------------------------------------------------------------------
// Exact face flux (ausm.js:183-393)
// ---------------------------------------------------------------------------
// The limiter functions take the MUSCL ratio r below, which is +0 or
// 2^-72 <= |r| <= 2^50 for any finite state (numerator +0 or at least
// 2^-44 in magnitude, divisor in [1e-6, ~1e8]): the numerators r^2 + r and
// r + |r| are then +0 or at least 2^-72 in magnitude (r^2 + r near r = -1
// is a multiple of 2^-48) and the divisors lie in [1, 2^101], so div_rn is
// the IEEE quotient.
template <typename R>
__device__ __forceinline__ R van_albada_exact(R r) {
  const R r2 = r * r;
  return div_rn(r2 + r, r2 + R(1.0));
}
template <typename R>
__device__ __forceinline__ R van_leer_exact(R r) {
  const R a = fabs(r);
  return div_rn(r + a, R(1.0) + a);
}
template <typename R>
__device__ __forceinline__ R minmod_exact(R a, R b) {
  return (R(0.5) * (wsign(a) + wsign(b))) * fmin(fabs(a), fabs(b));
}

// MUSCL on one primitive component (ausm.js:240-249)
// component-wise a - b (the stencil differences MUSCL works on)
template <typename R>
__device__ __forceinline__ PrimT<R> pdiff(PrimT<R> a, PrimT<R> b) {
  return PrimT<R>{a.r - b.r, a.u - b.u, a.v - b.v, a.p - b.p};
}

// Limiter freezing, a build extension with no counterpart in the reference
// (a convergence device for steady solves).  With r_L = (dL dF + eps) /
// (dF^2 + eps), r_R = (dF dR + eps) / (dF^2 + eps), psi = phi(r) / r and
// w = dF^2 / (dF^2 + eps), MUSCL's limited face values are exactly
//   L = W_L + psi_L (w dL + (1 - w) dF) / 2,
//   R = W_R - psi_R (w dR + (1 - w) dF) / 2
// (phi(r) dF = psi r dF and r dF = w dL + (1 - w) dF).  Recording takes
// psi = (r + 1) / (r^2 + 1) of van Albada (0 for r <= 0 with the clamped
// form) and w of every face and primitive component from the current state,
// in IEEE arithmetic in both math modes; a frozen face reconstructs with
// the recorded psi and w, an operator that equals the limited one at the
// recorded state and no longer switches.  Per face 12 values: psi_L of
// (rho, u, v, p), psi_R, w.  mode 0: normal; 1: normal, and the 12 values
// into rec; 2: from psi.
template <typename R>
struct FreezeT {
  int mode;
  const R* psi;  // 12 values (mode 2)
  R* rec;        // 12 values (mode 1)
};
template <typename R>
__device__ __forceinline__ R psi_of(int limiter, R r) {
  if (limiter == LIM_VANALBADA_CLAMPED && !(r > R(0))) return R(0);
  return (r + R(1)) / (r * r + R(1));
}
// the limiter ratios of the reference's MUSCL (ausm.js:244-245), IEEE, and
// the weight w of the upwind slope
template <typename R>
__device__ __forceinline__ void muscl_ratios(R dL, R dF, R dR, R& rL, R& rR, R& w) {
  const R eps = R(1e-6);
  rL = (dL * dF + eps) / (dF * dF + eps);
  rR = (dF * dR + eps) / (dF * dF + eps);
  w = (dF * dF) / (dF * dF + eps);
}

// on the stencil differences dL = wl - wl2, dF = wr - wl, dR = wr2 - wr,
// which a caller may share between neighbouring faces
template <typename R>
__device__ __forceinline__ void muscl_exact(int limiter, R wl, R wr, R dL, R dF, R dR, R& L,
                                            R& Rr) {
  const R eps = R(1e-6);
  if (limiter == LIM_MINMOD) {
    L = wl + R(0.5) * minmod_exact(dL, dF);
    Rr = wr - R(0.5) * minmod_exact(dF, dR);
  } else {
    // the ratios share their divisor dF^2 + eps, at least 1e-6 and at most
    // ~1e8 for primitive differences of any physical state; each numerator
    // x + eps is +0 (exact cancellation) or at least 2^-44 in magnitude
    // (nearly cancelling, x + eps is a multiple of the smaller ulp of its
    // two terms; otherwise it is near eps or near x), so div_by is the
    // IEEE quotient
    const R den = dF * dF + eps;
    const R yd = recip_rn(den);
    const R rL = div_by(dL * dF + eps, den, yd);
    const R rR = div_by(dF * dR + eps, den, yd);
    R phiL = limiter == LIM_VANLEER ? van_leer_exact(rL) : van_albada_exact(rL);
    R phiR = limiter == LIM_VANLEER ? van_leer_exact(rR) : van_albada_exact(rR);
    if (limiter == LIM_VANALBADA_CLAMPED) {
      // van Albada as a TVD limiter: no reconstruction where r <= 0
      phiL = rL > R(0) ? phiL : R(0);
      phiR = rR > R(0) ? phiR : R(0);
    }
    L = wl + (R(0.5) * phiL) * dF;
    Rr = wr - (R(0.5) * phiR) * dF;
  }
}

// psi and w of the four components (limiter freezing, record)
template <typename R>
__device__ __forceinline__ void record_psi(int limiter, PrimT<R> dL, PrimT<R> dF, PrimT<R> dR,
                                           R* rec) {
  const R l[4] = {dL.r, dL.u, dL.v, dL.p}, f[4] = {dF.r, dF.u, dF.v, dF.p},
          g[4] = {dR.r, dR.u, dR.v, dR.p};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    R rL, rR, w;
    muscl_ratios(l[k], f[k], g[k], rL, rR, w);
    rec[k] = psi_of(limiter, rL);
    rec[4 + k] = psi_of(limiter, rR);
    rec[8 + k] = w;
  }
}

// b = W_L, c = W_R and the stencil differences dL = W_L - W_L2,
// dF = W_R - W_L, dR = W_R2 - W_R
template <int SCHEME, typename Real>
__device__ __forceinline__ FluxT<Real> face_flux_exact(const DevParamsT<Real>& P, int limiter, PrimT<Real> b,
                                                   PrimT<Real> c, PrimT<Real> dL, PrimT<Real> dF,
                                                   PrimT<Real> dR, Real nx, Real ny,
                                                   FreezeT<Real> frz = FreezeT<Real>{0, nullptr, nullptr}) {
  const Real gamma = P.gamma;
  const Real gammaMinus1 = gamma - Real(1.0);
  PrimT<Real> L = b, R = c;
  if (limiter != LIM_NONE && frz.mode == 2) {
    const Real* s = frz.psi;
    auto fl = [&](Real w0, Real psL, Real psR, Real d0, Real df, Real d1, Real& l, Real& r) {
      l = l + (Real(0.5) * psL) * (w0 * d0 + (Real(1) - w0) * df);
      r = r - (Real(0.5) * psR) * (w0 * d1 + (Real(1) - w0) * df);
    };
    fl(s[8], s[0], s[4], dL.r, dF.r, dR.r, L.r, R.r);
    fl(s[9], s[1], s[5], dL.u, dF.u, dR.u, L.u, R.u);
    fl(s[10], s[2], s[6], dL.v, dF.v, dR.v, L.v, R.v);
    fl(s[11], s[3], s[7], dL.p, dF.p, dR.p, L.p, R.p);
  } else if (limiter != LIM_NONE) {
    muscl_exact(limiter, b.r, c.r, dL.r, dF.r, dR.r, L.r, R.r);
    muscl_exact(limiter, b.u, c.u, dL.u, dF.u, dR.u, L.u, R.u);
    muscl_exact(limiter, b.v, c.v, dL.v, dF.v, dR.v, L.v, R.v);
    muscl_exact(limiter, b.p, c.p, dL.p, dF.p, dR.p, L.p, R.p);
    if (frz.mode == 1) record_psi(limiter, dL, dF, dR, frz.rec);
  }
  const Real u_L = L.u * nx + L.v * ny;
  const Real u_R = R.u * nx + R.v * ny;
  const Real vvL = L.u * L.u + L.v * L.v;
  const Real vvR = R.u * R.u + R.v * R.v;
  // p / ((gamma-1) rho) and p / rho: the reconstructed pressure is
  // W.p >= 1e-7 plus a rounded product, so it is +0 or at least ~1e-15 in
  // magnitude, and rho of a physical state is far inside div_rn's range
  const Real h_L = (div_rn(L.p, gammaMinus1 * L.r) + Real(0.5) * vvL) + div_rn(L.p, L.r);
  const Real h_R = (div_rn(R.p, gammaMinus1 * R.r) + Real(0.5) * vvR) + div_rn(R.p, R.r);
  Real mdot, P_i;
  if constexpr (SCHEME == SCHEME_AUSMUP) {
    // ausm.js:273-327
    // free-stream numerators (|V|^2 is +0 or a normal number, gamma p_inf > 0)
    const Real M_inf2 = div_rn(P.vx * P.vx + P.vy * P.vy, div_rn(gamma * P.inP, L.r));
    const Real sigma = Real(1.0);
    const Real h_t = Real(0.5) * (h_L + h_R);
    // argument at least 1e-6
    const Real aStar = sqrt_rn(fmax(Real(1e-6), ((Real(2.0) * gammaMinus1) / (gamma + Real(1.0))) * h_t));
    // aStar >= 1e-3, and the divisors are at least aStar
    const Real aHat_L = div_rn(aStar * aStar, fmax(u_L, aStar));
    const Real aHat_R = div_rn(aStar * aStar, fmax(-u_R, aStar));
    const Real a_i = fmin(aHat_L, aHat_R);
    // a_i >= a*^2 / max(|u|, a*) is a sound speed of the state: inside
    // div_free's divisor range, and so are 2 a_i^2 and rho_i a_i^2 below
    const Real un[2] = {u_L, u_R};
    Real Mn[2];
    div_free(un, a_i, Mn);
    const Real M_L = Mn[0];
    const Real M_R = Mn[1];
    const Real a_i2 = a_i * a_i;
    const Real mb[1] = {u_L * u_L + u_R * u_R};
    Real mq[1];
    div_free(mb, a_i2 * Real(2.0), mq);
    const Real Mbar2 = mq[0];
    const Real M_o2 = fmin(Real(1.0), fmax(Mbar2, M_inf2));
    const Real f_a = M_o2 * (Real(2.0) - M_o2);
    const Real rho_i = Real(0.5) * (L.r + R.r);
    const Real Mp1L = M_L + Real(1.0), Mp1R = M_R + Real(1.0);
    const Real Mm1L = M_L - Real(1.0), Mm1R = M_R - Real(1.0);
    const Real M2pL = (Real(0.25) * Mp1L) * Mp1L, M2pR = (Real(0.25) * Mp1R) * Mp1R;
    const Real M2mL = (-Real(0.25) * Mm1L) * Mm1L, M2mR = (-Real(0.25) * Mm1R) * Mm1R;
    const Real M4subL = M2pL * (Real(1.0) + Real(2.0) * M2mL);
    const Real M4subR = M2mR * (Real(1.0) - Real(2.0) * M2pR);
    const Real supL = wstep(Real(1.0), fabs(M_L)), supR = wstep(Real(1.0), fabs(M_R));
    const Real M4L = wsel(M4subL, fmax(Real(0.0), M_L), supL);
    const Real M4R = wsel(M4subR, fmin(Real(0.0), M_R), supR);
    Real kp = P.Kp / f_a;
    kp = kp * fmax(Real(1.0) - sigma * Mbar2, Real(0.0));
    kp = kp * (R.p - L.p);
    const Real kn[1] = {kp};
    Real kq[1];
    div_free(kn, rho_i * a_i2, kq);
    kp = kq[0];
    const Real M_i = (M4L + M4R) - kp;
    const Real alpha4 = Real(0.75) * (-Real(4.0) + (Real(5.0) * f_a) * f_a);
    const Real P5subL = M2pL * ((Real(2.0) - M_L) + (alpha4 * M_L) * M2mL);
    const Real P5subR = M2mR * ((-Real(2.0) - M_R) - (alpha4 * M_R) * M2pR);
    const Real P5L = wsel(P5subL, wstep(Real(0.0), M_L), supL);
    const Real P5R = wsel(P5subR, wstep(Real(0.0), -M_R), supR);
    Real ku = P.Ku * P5L;
    ku = ku * P5R;
    ku = ku * (L.r + R.r);
    ku = ku * f_a;
    ku = ku * a_i;
    ku = ku * (u_R - u_L);
    P_i = (P5L * L.p + P5R * R.p) - ku;
    mdot = (a_i * M_i) * (M_i > Real(0.0) ? L.r : R.r);
  } else {
    // ausm.js:348-384
    // numerators at least 1.4e-10, divisors at least 1e-10
    // square roots of at least 1.4e-10 / rho: in sqrt_rn's range for any
    // density below 2^60
    const Real a_L = sqrt_rn(div_rn(gamma * fmax(L.p, Real(1e-10)), fmax(L.r, Real(1e-10))));
    const Real a_R = sqrt_rn(div_rn(gamma * fmax(R.p, Real(1e-10)), fmax(R.r, Real(1e-10))));
    const Real a_i = (a_L + a_R) * Real(0.5);
    const Real Uarc = sqrt(Real(0.5) * (vvL + vvR));
    // a_i, the mean sound speed, is inside div_free's divisor range
    const Real sn[3] = {u_L, u_R, Uarc};
    Real sq[3];
    div_free(sn, a_i, sq);
    const Real M_L = sq[0];
    const Real M_R = sq[1];
    const Real Marc = fmin(Real(1.0), sq[2]);
    const Real X = (Real(1.0) - Marc) * (Real(1.0) - Marc);
    const Real qL = M_L + Real(1.0), qR = M_R - Real(1.0);
    const Real sqL = (Real(0.5) * qL) * qL, sqR = (Real(0.5) * qR) * qR;
    const Real supL = wstep(Real(1.0), fabs(M_L)), supR = wstep(Real(1.0), fabs(M_R));
    const Real fpp = wsat(Real(0.5) * wsel(sqL * (Real(2.0) - Real(1.0) * M_L), Real(1.0) + Real(1.0) * wsign(M_L), supL));
    const Real fpm =
        wsat(Real(0.5) * wsel(sqR * (Real(2.0) - -Real(1.0) * M_R), Real(1.0) + -Real(1.0) * wsign(M_R), supR));
    if constexpr (SCHEME == SCHEME_SLAU2) {
      const Real t1 = (fpp - fpm) * (L.p - R.p);
      const Real t2 = ((Uarc * ((fpp + fpm) - Real(1.0))) * (L.r + R.r)) * a_i;
      P_i = Real(0.5) * ((t1 + (L.p + R.p)) + t2);
    } else {
      const Real t1 = (fpp - fpm) * (L.p - R.p);
      const Real t2 = (Real(1.0) + (Real(1.0) - X) * ((fpp + fpm) - Real(1.0))) * (L.p + R.p);
      P_i = Real(0.5) * (t1 + t2);
    }
    const Real delta_p = R.p - L.p;
    const Real ubn[1] = {L.r * fabs(u_L) + R.r * fabs(u_R)};
    Real ubq[1];
    div_free(ubn, L.r + R.r, ubq);
    const Real ubar_n = ubq[0];
    const Real g = wsat(-M_L) * wsat(M_R);
    const Real ubp = wmix(ubar_n, fabs(u_L), g);
    const Real ubm = wmix(ubar_n, fabs(u_R), g);
    // X = (1 - Marc)^2 with Marc in [0, 1] is +0 or at least 2^-48
    mdot = Real(0.5) * ((L.r * (u_L + ubp) + R.r * (u_R - ubm)) * (Real(1.0) - g) - div_rn(X, a_i) * delta_p);
  }
  const bool left = mdot > Real(0.0);
  FluxT<Real> F;
  F.f0 = Real(1.0) * mdot + Real(0.0);
  F.f1 = (left ? L.u : R.u) * mdot + nx * P_i;
  F.f2 = (left ? L.v : R.v) * mdot + ny * P_i;
  F.f3 = (left ? h_L : h_R) * mdot + Real(0.0);
  return F;
}

// ---------------------------------------------------------------------------
// Fast face flux: the same scheme with shared reciprocals and contraction.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void muscl_fast(int limiter, float wl, float wr, float dL, float dF, float dR,
                                           float& L, float& Rr) {
  const float eps = 1e-6f;
  const float hF = 0.5f * dF;
  if (limiter == LIM_MINMOD) {
    const float mL = (dL * dF > 0.0f) ? copysignf(fminf(fabsf(dL), fabsf(dF)), dF) : 0.0f;
    const float mR = (dF * dR > 0.0f) ? copysignf(fminf(fabsf(dF), fabsf(dR)), dF) : 0.0f;
    L = fmaf(0.5f, mL, wl);
    Rr = fmaf(-0.5f, mR, wr);
  } else if (limiter == LIM_VANLEER) {
    // phi(r) = (r+|r|)/(1+|r|) with r = n/d, d > 0:  (n+|n|)/(d+|n|)
    const float d = fmaf(dF, dF, eps);
    const float nL = fmaf(dL, dF, eps), nR = fmaf(dF, dR, eps);
    const float phiL = (nL + fabsf(nL)) * __builtin_amdgcn_rcpf(d + fabsf(nL));
    const float phiR = (nR + fabsf(nR)) * __builtin_amdgcn_rcpf(d + fabsf(nR));
    L = fmaf(phiL, hF, wl);
    Rr = fmaf(-phiR, hF, wr);
  } else {
    // van Albada phi(r) = (r^2+r)/(r^2+1) with r = n/d:  n(n+d)/(n^2+d^2);
    // one reciprocal serves both sides: 1/den_L = den_R / (den_L den_R).
    // den >= d^2 >= eps^2 keeps the product normal; if it overflows the
    // reciprocal is 0 and both sides fall back to first order.
    // The half difference dF/2 rides on the shared reciprocal:
    // phi_L dF/2 = n_L (n_L + d) * (den_R * (dF/2) / (den_L den_R)).
    const float d = fmaf(dF, dF, eps);
    const float d2 = d * d;
    const float nL = fmaf(dL, dF, eps), nR = fmaf(dF, dR, eps);
    const float denL = fmaf(nL, nL, d2), denR = fmaf(nR, nR, d2);
    const float rh = __builtin_amdgcn_rcpf(denL * denR) * hF;
    float aL = nL * (nL + d), aR = nR * (nR + d);
    if (limiter == LIM_VANALBADA_CLAMPED) {
      // r > 0 exactly when n > 0 (d > 0): the clamped limiter's phi = 0
      aL = nL > 0.0f ? aL : 0.0f;
      aR = nR > 0.0f ? aR : 0.0f;
    }
    L = fmaf(aL, denR * rh, wl);
    Rr = fmaf(-aR, denL * rh, wr);
  }
}

// Returns the flux already multiplied by the face length len; (sx, sy) =
// len * (nx, ny) is the unnormalised edge normal, which carries the pressure
// term without a separate scaling.
template <int SCHEME>
__device__ __forceinline__ Flux face_flux_fast(const DevParams& P, int limiter, Prim b, Prim c,
                                               Prim dL, Prim dF, Prim dR, float nx, float ny,
                                               float len, float sx, float sy,
                                               FreezeT<float> frz = FreezeT<float>{0, nullptr, nullptr}) {
  const float gamma = P.gamma;
  Prim L = b, R = c;
  if (limiter != LIM_NONE && frz.mode == 2) {
    // w d + (1 - w) dF = dF + w (d - dF)
    const float* s = frz.psi;
    auto fl = [&](float w0, float psL, float psR, float d0, float df, float d1, float& l, float& r) {
      l = fmaf(0.5f * psL, fmaf(w0, d0 - df, df), l);
      r = fmaf(-0.5f * psR, fmaf(w0, d1 - df, df), r);
    };
    fl(s[8], s[0], s[4], dL.r, dF.r, dR.r, L.r, R.r);
    fl(s[9], s[1], s[5], dL.u, dF.u, dR.u, L.u, R.u);
    fl(s[10], s[2], s[6], dL.v, dF.v, dR.v, L.v, R.v);
    fl(s[11], s[3], s[7], dL.p, dF.p, dR.p, L.p, R.p);
  } else if (limiter != LIM_NONE) {
    muscl_fast(limiter, b.r, c.r, dL.r, dF.r, dR.r, L.r, R.r);
    muscl_fast(limiter, b.u, c.u, dL.u, dF.u, dR.u, L.u, R.u);
    muscl_fast(limiter, b.v, c.v, dL.v, dF.v, dR.v, L.v, R.v);
    muscl_fast(limiter, b.p, c.p, dL.p, dF.p, dR.p, L.p, R.p);
    if (frz.mode == 1) record_psi(limiter, dL, dF, dR, frz.rec);
  }
  const float u_L = fmaf(L.u, nx, L.v * ny);
  const float u_R = fmaf(R.u, nx, R.v * ny);
  const float vvL = fmaf(L.u, L.u, L.v * L.v);
  const float vvR = fmaf(R.u, R.u, R.v * R.v);
  // 1/rho_L, 1/rho_R and 1/(rho_L + rho_R) from one reciprocal
  const float rsum = L.r + R.r;
  const float rprod = L.r * R.r;
  const float rall = __builtin_amdgcn_rcpf(rprod * rsum);
  const float irS = rprod * rall;
  const float irL = (R.r * rsum) * rall, irR = (L.r * rsum) * rall;
  // 1 / max(rho, 1e-10), the sound-speed guard
  const float iclL = L.r > 1e-10f ? irL : 1e10f, iclR = R.r > 1e-10f ? irR : 1e10f;
  // h = p / ((gamma-1) rho) + |v|^2 / 2 + p / rho = gamma/(gamma-1) p/rho + |v|^2 / 2
  const float h_L = fmaf(P.ggm1, L.p * irL, 0.5f * vvL);
  const float h_R = fmaf(P.ggm1, R.p * irR, 0.5f * vvR);
  float mdot, P_i;
  if constexpr (SCHEME == SCHEME_AUSMUP) {
    const float M_inf2 = P.minf_coef * L.r;
    // astar_coef * (h_L + h_R)/2, with the exact halving moved onto the
    // (uniform) coefficient
    const float aStar = __builtin_amdgcn_sqrtf(fmaxf(1e-6f, (0.5f * P.astar_coef) * (h_L + h_R)));
    const float as2 = aStar * aStar;
    // min(a*^2 / max(u_L, a*), a*^2 / max(-u_R, a*)) = a*^2 / max(u_L, -u_R, a*);
    // a_i and 1/a_i from one reciprocal
    const float mx = fmaxf(fmaxf(u_L, -u_R), aStar);
    const float rq = __builtin_amdgcn_rcpf(mx * as2);
    const float a_i = (as2 * as2) * rq;
    const float ia = (mx * mx) * rq;
    const float M_L = u_L * ia;
    const float M_R = u_R * ia;
    const float a_i2 = a_i * a_i;
    const float Mbar2 = 0.5f * fmaf(M_L, M_L, M_R * M_R);
    const float M_o2 = fminf(1.0f, fmaxf(Mbar2, M_inf2));
    const float f_a = M_o2 * (2.0f - M_o2);
    const float rho_i = 0.5f * (L.r + R.r);
    // M2+(M) = ((M + 1)/2)^2, M2-(M) = -((M - 1)/2)^2
    const float hpL = fmaf(0.5f, M_L, 0.5f), hpR = fmaf(0.5f, M_R, 0.5f);
    const float hmL = fmaf(0.5f, M_L, -0.5f), hmR = fmaf(0.5f, M_R, -0.5f);
    const float M2pL = hpL * hpL, M2pR = hpR * hpR;
    const float M2mL = -(hmL * hmL), M2mR = -(hmR * hmR);
    const bool supL = fabsf(M_L) >= 1.0f, supR = fabsf(M_R) >= 1.0f;
    const float M4L = supL ? fmaxf(0.0f, M_L) : M2pL * fmaf(2.0f, M2mL, 1.0f);
    const float M4R = supR ? fminf(0.0f, M_R) : M2mR * fmaf(-2.0f, M2pR, 1.0f);
    float M_i = M4L + M4R;
    if (P.Kp != 0.0f) {
      const float kp = P.Kp * fmaxf(1.0f - Mbar2, 0.0f) * (R.p - L.p) *
                       __builtin_amdgcn_rcpf(f_a * rho_i * a_i2);
      M_i -= kp;
    }
    const float alpha4 = fmaf(3.75f * f_a, f_a, -3.0f);  // 3/4 (5 f_a^2 - 4)
    // both sides evaluated, then selected
    const float p5sL = M2pL * fmaf(alpha4 * M_L, M2mL, 2.0f - M_L);
    const float p5sR = M2mR * fmaf(-alpha4 * M_R, M2pR, -2.0f - M_R);
    const float stL = M_L >= 0.0f ? 1.0f : 0.0f, stR = -M_R >= 0.0f ? 1.0f : 0.0f;
    const float P5L = supL ? stL : p5sL;
    const float P5R = supR ? stR : p5sR;
    const float ku = P.Ku * P5L * P5R * (L.r + R.r) * f_a * a_i * (u_R - u_L);
    P_i = fmaf(P5L, L.p, fmaf(P5R, R.p, -ku));
    mdot = a_i * M_i * (M_i > 0.0f ? L.r : R.r);
  } else {
    const float a_L = __builtin_amdgcn_sqrtf(gamma * fmaxf(L.p, 1e-10f) * iclL);
    const float a_R = __builtin_amdgcn_sqrtf(gamma * fmaxf(R.p, 1e-10f) * iclR);
    const float a_i = (a_L + a_R) * 0.5f;
    const float ia = __builtin_amdgcn_rcpf(a_i);
    const float M_L = u_L * ia;
    const float M_R = u_R * ia;
    const float Uarc = __builtin_amdgcn_sqrtf(0.5f * (vvL + vvR));
    const float Marc = fminf(1.0f, Uarc * ia);
    const float X = (1.0f - Marc) * (1.0f - Marc);
    const bool supL = fabsf(M_L) >= 1.0f, supR = fabsf(M_R) >= 1.0f;
    const float qL = M_L + 1.0f, qR = M_R - 1.0f;
    // supersonic side: |M| >= 1, so (1 +- sign(M))/2 is exactly the step M >< 0
    const float fpp = wsat(supL ? (M_L > 0.0f ? 1.0f : 0.0f) : 0.25f * qL * qL * (2.0f - M_L));
    const float fpm = wsat(supR ? (M_R < 0.0f ? 1.0f : 0.0f) : 0.25f * qR * qR * (2.0f + M_R));
    const float fsum = (fpp + fpm) - 1.0f;
    if constexpr (SCHEME == SCHEME_SLAU2) {
      P_i = 0.5f * fmaf(fpp - fpm, L.p - R.p, fmaf(Uarc * fsum, (L.r + R.r) * a_i, L.p + R.p));
    } else {
      P_i = 0.5f * fmaf(fpp - fpm, L.p - R.p, fmaf(1.0f - X, fsum, 1.0f) * (L.p + R.p));
    }
    const float delta_p = R.p - L.p;
    const float ubar_n = fmaf(L.r, fabsf(u_L), R.r * fabsf(u_R)) * irS;
    const float g = wsat(-M_L) * wsat(M_R);
    const float ubp = fmaf(g, fabsf(u_L) - ubar_n, ubar_n);
    const float ubm = fmaf(g, fabsf(u_R) - ubar_n, ubar_n);
    const float conv = fmaf(L.r, u_L + ubp, R.r * (u_R - ubm));
    mdot = 0.5f * fmaf(conv, 1.0f - g, -(X * ia) * delta_p);
  }
  const bool left = mdot > 0.0f;
  const float ml = mdot * len;
  Flux F;
  F.f0 = ml;
  F.f1 = fmaf(left ? L.u : R.u, ml, sx * P_i);
  F.f2 = fmaf(left ? L.v : R.v, ml, sy * P_i);
  F.f3 = (left ? h_L : h_R) * ml;
  return F;
}

// ---------------------------------------------------------------------------
// Ghost states (boundary pass, ausm.js:34-120).  Computed on the fly from the
// first / last interior row instead of being stored.
// ---------------------------------------------------------------------------

// division and square root of the ghost-row formulas: correctly rounded, or
// (FAST, FP32 only) the hardware reciprocal and square root
template <bool FAST, typename R>
__device__ __forceinline__ R gdiv(R a, R b) {
  if constexpr (FAST) return a * __builtin_amdgcn_rcpf(b);
  else return a / b;
}
template <bool FAST, typename R>
__device__ __forceinline__ R gsqrt(R a) {
  if constexpr (FAST) return __builtin_amdgcn_sqrtf(a);
  else return sqrt(a);
}

// slip wall: reflect the velocity of interior row 1 about the body-segment
// normal (ausm.js:57-64).  (nx, ny) = rot90(normalize(v_i - v_{i+1})).
template <typename R, bool FAST = false>
__device__ __forceinline__ ConsT<R> wall_ghost(ConsT<R> I, R nx, R ny) {
  const R rho = I.r;
  const R ux = gdiv<FAST>(I.mu, rho), uy = gdiv<FAST>(I.mv, rho);
  const R k = R(2.0) * (nx * ux + ny * uy);
  const R gx = ux - k * nx, gy = uy - k * ny;
  ConsT<R> g;
  g.r = rho;
  g.mu = gx * rho;
  g.mv = gy * rho;
  g.e = I.e;
  return g;
}

// far field from interior row N (ausm.js:65-117).  second = the outer ghost
// row (N+2), which extrapolates linearly on outflow.
template <typename R, bool FAST = false>
__device__ __forceinline__ ConsT<R> far_ghost(const DevParamsT<R>& P, ConsT<R> I, R nx, R ny,
                                          bool second) {
  const R gamma = P.gamma;
  const R rho = I.r;
  const R ux = gdiv<FAST>(I.mu, rho), uy = gdiv<FAST>(I.mv, rho);
  const R bio = nx * P.vx + ny * P.vy;
  ConsT<R> g;
  if (bio < R(0.0)) {
    g.r = P.inState[0];
    g.mu = P.inState[1];
    g.mv = P.inState[2];
    g.e = P.inState[3];
    return g;
  }
  const R gm1inv = gdiv<FAST>(R(1.0), gamma - R(1.0));
  const R u2 = ux * ux + uy * uy;
  const R Pr = (I.e - (R(0.5) * u2) * rho) * (gamma - R(1.0));
  const R a2 = gdiv<FAST>(gamma * Pr, rho);
  const R a2inf = gdiv<FAST>(gamma * P.inP, P.inRho);
  const R M_int = gdiv<FAST>(u2, a2);
  const R M_inf2 = gdiv<FAST>(P.vx * P.vx + P.vy * P.vy, a2inf);
  if (M_int < R(1.0) && M_inf2 < R(1.0)) {
    const R un_int = ux * nx + uy * ny;
    const R un_inf = bio;
    const R a_int = gsqrt<FAST>(a2);
    const R a_inf = gsqrt<FAST>(a2inf);
    const R J_int = un_int + (R(2.0) * a_int) * gm1inv;
    const R J_ext = un_inf - (R(2.0) * a_inf) * gm1inv;
    const R un_g = R(0.5) * (J_int + J_ext);
    const R a_g = fmax(a_int * R(0.1), (R(0.25) * (J_int - J_ext)) * (gamma - R(1.0)));
    const R X = fmax(R(1e-6), gdiv<FAST>(a_g, a_int));
    const R X2 = X * X;
    const R rho_g = fmax(R(1e-6), ((rho * X2) * X2) * X);
    const R dn = un_g - un_int;
    const R vgx = dn * nx + ux, vgy = dn * ny + uy;
    const R rhoE_g = fmax(R(1e-6), gdiv<FAST>((rho_g * a_g) * a_g, gamma)) * gm1inv +
                         (R(0.5) * rho_g) * (vgx * vgx + vgy * vgy);
    g.r = rho_g;
    g.mu = vgx * rho_g;
    g.mv = vgy * rho_g;
    g.e = rhoE_g;
  } else if (M_inf2 > R(1.0)) {
    g.r = I.r;
    g.mu = I.mu;
    g.mv = I.mv;
    g.e = gdiv<FAST>(P.inP, gamma - R(1.0)) + (R(0.5) * rho) * u2;
  } else {
    g = I;
  }
  if (second) {
    g.r = R(2.0) * g.r - I.r;
    g.mu = R(2.0) * g.mu - I.mu;
    g.mv = R(2.0) * g.mv - I.mv;
    g.e = R(2.0) * g.e - I.e;
  }
  return g;
}

// the outer far-field ghost (row N+2) from the inner one (row N+1): the
// same state on inflow, 2 g - I on outflow (ausm.js:114-116) -- what
// far_ghost(..., true) returns, without evaluating the boundary state again
template <typename R>
__device__ __forceinline__ ConsT<R> far_ghost_outer(const DevParamsT<R>& P, ConsT<R> g, ConsT<R> I,
                                                R nx, R ny) {
  const R bio = nx * P.vx + ny * P.vy;
  if (bio < R(0.0)) return g;
  return ConsT<R>{R(2.0) * g.r - I.r, R(2.0) * g.mu - I.mu, R(2.0) * g.mv - I.mv,
                  R(2.0) * g.e - I.e};
}

}  // namespace ausm
