/*
 * CPU ORACLE — test infrastructure only.
 *
 * A plain-C restatement of the reference's explicit O-grid solver
 * (huj31415/ausm-webgpu), written from the WGSL text as the specification and
 * kept in the reference's own unfused structure: per RK stage a boundary
 * pass, a vertical-flux pass, a horizontal-flux pass, a residual pass and a
 * stage update (15 passes per step), with the same three state textures, the
 * same buffer rotation and the same lagged CFL time step.
 *
 * This file is compiled twice:
 *   -DREAL=float  -DPFX=orc32_   ("ref32": FP32, no FMA contraction)
 *   -DREAL=double -DPFX=orc64_   ("ref64": the same scheme in FP64)
 * Only tests/, the smoke check and bench.py's CPU-baseline leg load it, and
 * only as the checker; the product (the HIP library) never links it.
 *
 * Layouts follow the reference textures, row-major with x (the i index around
 * the body) fastest, AoS per texel:
 *   vtx     (N+1) x M x 2   gridPoints   (main.js:110)
 *   state   (N+3) x M x 4   state0/1/2   (main.js:117-119); row 0 body ghost,
 *                           rows 1..N interior, rows N+1, N+2 far-field ghosts
 *   fluxX   N x (M+1) x 4   (main.js:121)  column M is never read
 *   fluxY   (N+1) x M x 4   (main.js:122)
 *   res     N x M x 4       (main.js:123)
 *   area    N x M           (main.js:113)
 *   flen    N x M x 4       (left, right, up, down)  (main.js:114)
 *
 * WGSL built-ins are restated with these fixed definitions (the WGSL spec
 * leaves some implementation-defined; these choices are the ones the HIP
 * kernels implement as well):
 *   dot(a,b)      = a.x*b.x + a.y*b.y (+ ...), summed left to right
 *   length(v)     = sqrt(dot(v,v));  normalize(v) = v / length(v)
 *   mix(a,b,t)    = a*(1-t) + b*t
 *   reflect(e,n)  = e - (2*dot(n,e))*n
 *   step(e,x)     = x >= e ? 1 : 0;  sign(0) = 0;  saturate = clamp(x,0,1)
 *   min/max       = C fmin/fmax
 *
 * Parity status of each function is recorded in the header of tests that pin
 * it; functions the tests leave open are marked "parity unpinned" below.
 */
#pragma STDC FP_CONTRACT OFF
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#error "compile with -DREAL=float or -DREAL=double"
#endif
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(PFX, name)

typedef REAL R;

/* parameter vector layout (mirrors the Uniforms fields used by the solver,
 * util.js:2-27) */
enum {
  P_GAMMA = 0, P_INRHO = 1, P_INP = 2, P_VX = 3, P_VY = 4,
  P_IS0 = 5, P_IS1 = 6, P_IS2 = 7, P_IS3 = 8,
  P_MAXDT = 9, P_CFL = 10, P_KP = 11, P_KU = 12, P_COUNT = 13
};
enum { SCHEME_SLAU2 = 0, SCHEME_SLAU = 1, SCHEME_AUSMUP = 2 }; /* util.js:114-118 */
enum { LIM_VANALBADA = 0, LIM_MINMOD = 1, LIM_VANLEER = 2, LIM_NONE = 3, LIM_VANALBADA_CLAMPED = 4 };

static inline R rmin(R a, R b) { return sizeof(R) == 4 ? (R)fminf((float)a, (float)b) : (R)fmin(a, b); }
static inline R rmax(R a, R b) { return sizeof(R) == 4 ? (R)fmaxf((float)a, (float)b) : (R)fmax(a, b); }
static inline R rsqrt_(R a) { return sizeof(R) == 4 ? (R)sqrtf((float)a) : (R)sqrt(a); }
static inline R rabs(R a) { return sizeof(R) == 4 ? (R)fabsf((float)a) : (R)fabs(a); }
static inline R rsat(R x) { return rmin(rmax(x, (R)0), (R)1); }
static inline R rsign(R x) { return x > (R)0 ? (R)1 : (x < (R)0 ? (R)-1 : (R)0); }
static inline R rstep(R edge, R x) { return x >= edge ? (R)1 : (R)0; }
static inline R rmix(R a, R b, R t) { return a * ((R)1 - t) + b * t; }

/* ------------------------------------------------------------------------ */
/* state helpers: ausm.js:138-157                                            */
/* ------------------------------------------------------------------------ */

/* loadState: x wraps mod M, y is shifted by one ghost row and clamped at the
 * body ghost (ausm.js:138-143). */
static inline const R* load_state(const R* st, int M, int x, int y) {
  int xi = (x + M) % M;
  int yi = y + 1;
  if (yi < 0) yi = 0;
  return st + ((size_t)yi * M + xi) * 4;
}

/* toPrimitive (ausm.js:145-150) */
static inline void to_primitive(const R* Q, R gamma, R* W) {
  R rho = Q[0];
  R vx = Q[1] / rho;
  R vy = Q[2] / rho;
  R d = vx * vx + vy * vy;
  R p = (Q[3] - ((R)0.5 * rho) * d) * (gamma - (R)1);
  W[0] = rho;
  W[1] = vx;
  W[2] = vy;
  W[3] = rmax((R)1e-7, p);
}

/* limiter functions (ausm.js:159-171) */
static inline R van_albada(R r) { R r2 = r * r; return (r2 + r) / (r2 + (R)1); }
static inline R van_leer(R r) { R a = rabs(r); return (r + a) / ((R)1 + a); }
static inline R minmod2(R a, R b) {
  return ((R)0.5 * (rsign(a) + rsign(b))) * rmin(rabs(a), rabs(b));
}

/* ------------------------------------------------------------------------ */
/* face flux: fluxInterfaceStateCode + scheme bodies (ausm.js:128-398)       */
/* ------------------------------------------------------------------------ */

/* Limiter freezing: a build extension with no counterpart in the reference
 * (a convergence device for steady MUSCL solves).  With the reference's
 * ratios r_L = (dL dF + eps) / (dF^2 + eps), r_R = (dF dR + eps) /
 * (dF^2 + eps), psi = phi(r)/r and w = dF^2 / (dF^2 + eps), the limited face
 * values are exactly L = W_L + psi_L (w dL + (1 - w) dF) / 2 and
 * R = W_R - psi_R (w dR + (1 - w) dF) / 2, since phi(r) dF = psi r dF and
 * r dF = w dL + (1 - w) dF.  Recording (mode 1) evaluates the limited flux
 * as usual and stores, per face and primitive component, psi = (r + 1) /
 * (r^2 + 1) of van Albada (0 for r <= 0 with the clamped form) and w; a
 * frozen face (mode 2) reconstructs with the stored values.  set_freeze
 * selects the mode and the record for every later flux pass: vertical faces
 * (N+1) x M x 12, then horizontal faces N x M x 12, each face's 12 values
 * psi_L of (rho, u, v, p), psi_R, w.  State of the library, not thread-safe:
 * callers set it around each call. */
static int FN(frz_mode) = 0;
static R* FN(frz_psi) = NULL;
void FN(set_freeze)(int mode, R* psi) {
  FN(frz_mode) = psi ? mode : 0;
  FN(frz_psi) = psi;
}
static inline R psi_of(int limiter, R r) {
  if (limiter == LIM_VANALBADA_CLAMPED && !(r > (R)0)) return (R)0;
  return (r + (R)1) / (r * r + (R)1);
}

/* One face flux from the four stencil states QL2, QL, QR, QR2 (conserved)
 * and the face unit normal n.  Writes the per-unit-length flux to F[4].
 * Follows ausm.js:183-267 (interface states), 271-336 (AUSM+-up) and
 * 345-393 (SLAU / SLAU2).  frz, psi_in, psi_out: limiter freezing (above). */
static void face_flux_frz(const R* par, int scheme, int limiter, const R* QL2, const R* QL,
                          const R* QR, const R* QR2, R nx, R ny, R* F, int frz,
                          const R* psi_in, R* psi_out) {
  const R gamma = par[P_GAMMA];
  const R gammaMinus1 = gamma - (R)1;
  const R eps = (R)1e-6;
  R WL[4], WL2[4], WR[4], WR2[4];
  to_primitive(QL, gamma, WL);
  to_primitive(QL2, gamma, WL2);
  to_primitive(QR, gamma, WR);
  to_primitive(QR2, gamma, WR2);

  R L[4], Rr[4];
  for (int c = 0; c < 4; ++c) { L[c] = WL[c]; Rr[c] = WR[c]; }
  if (limiter != LIM_NONE) {
    for (int c = 0; c < 4; ++c) {
      R dL = WL[c] - WL2[c];
      R dF = WR[c] - WL[c];
      R dR = WR2[c] - WR[c];
      if (frz == 2) {
        const R w = psi_in[8 + c];
        L[c] = L[c] + ((R)0.5 * psi_in[c]) * (w * dL + ((R)1 - w) * dF);
        Rr[c] = Rr[c] - ((R)0.5 * psi_in[4 + c]) * (w * dR + ((R)1 - w) * dF);
      } else if (limiter == LIM_MINMOD) {
        /* two-slope minmod (ausm.js:164-166), wired as a slope limiter */
        L[c] = L[c] + (R)0.5 * minmod2(dL, dF);
        Rr[c] = Rr[c] - (R)0.5 * minmod2(dF, dR);
      } else {
        R rL = (dL * dF + eps) / (dF * dF + eps);
        R rR = (dF * dR + eps) / (dF * dF + eps);
        R phiL = limiter == LIM_VANLEER ? van_leer(rL) : van_albada(rL);
        R phiR = limiter == LIM_VANLEER ? van_leer(rR) : van_albada(rR);
        if (limiter == LIM_VANALBADA_CLAMPED) {
          /* build extension (no counterpart in the reference): van Albada
           * as a TVD limiter, phi = 0 for r <= 0 */
          if (!(rL > (R)0)) phiL = (R)0;
          if (!(rR > (R)0)) phiR = (R)0;
        }
        L[c] = L[c] + ((R)0.5 * phiL) * dF;
        Rr[c] = Rr[c] - ((R)0.5 * phiR) * dF;
        if (frz == 1) {
          psi_out[c] = psi_of(limiter, rL);
          psi_out[4 + c] = psi_of(limiter, rR);
          psi_out[8 + c] = (dF * dF) / (dF * dF + eps);
        }
      }
    }
  }

  const R rho_L = L[0], vxL = L[1], vyL = L[2], p_L = L[3];
  const R rho_R = Rr[0], vxR = Rr[1], vyR = Rr[2], p_R = Rr[3];
  const R u_L = vxL * nx + vyL * ny;
  const R u_R = vxR * nx + vyR * ny;
  const R vvL = vxL * vxL + vyL * vyL;
  const R vvR = vxR * vxR + vyR * vyR;
  const R h_L = (p_L / (gammaMinus1 * rho_L) + (R)0.5 * vvL) + p_L / rho_L;
  const R h_R = (p_R / (gammaMinus1 * rho_R) + (R)0.5 * vvR) + p_R / rho_R;

  R mdot, P_interface;
  if (scheme == SCHEME_AUSMUP) {
    /* ausm.js:273-327 */
    const R vx = par[P_VX], vy = par[P_VY];
    const R M_inf2 = (vx * vx + vy * vy) / ((gamma * par[P_INP]) / rho_L);
    const R sigma = (R)1;
    const R h_t = (R)0.5 * (h_L + h_R);
    const R aStar = rsqrt_(rmax((R)1e-6, (((R)2 * gammaMinus1) / (gamma + (R)1)) * h_t));
    const R aHat_L = (aStar * aStar) / rmax(u_L, aStar);
    const R aHat_R = (aStar * aStar) / rmax(-u_R, aStar);
    const R a_i = rmin(aHat_L, aHat_R);
    const R M_L = u_L / a_i;
    const R M_R = u_R / a_i;
    const R a_i2 = a_i * a_i;
    const R Mbar2 = (u_L * u_L + u_R * u_R) / (a_i2 * (R)2);
    const R M_o2 = rmin((R)1, rmax(Mbar2, M_inf2));
    const R f_a = M_o2 * ((R)2 - M_o2);
    const R rho_i = (R)0.5 * (rho_L + rho_R);
    const R Mp1L = M_L + (R)1, Mp1R = M_R + (R)1;
    const R Mm1L = M_L - (R)1, Mm1R = M_R - (R)1;
    const R M2pL = ((R)0.25 * Mp1L) * Mp1L, M2pR = ((R)0.25 * Mp1R) * Mp1R;
    const R M2mL = ((R)-0.25 * Mm1L) * Mm1L, M2mR = ((R)-0.25 * Mm1R) * Mm1R;
    const R M4subL = M2pL * ((R)1 + (R)2 * M2mL);
    const R M4subR = M2mR * ((R)1 - (R)2 * M2pR);
    const R supL = rstep((R)1, rabs(M_L)), supR = rstep((R)1, rabs(M_R));
    const R M4L = rmix(M4subL, rmax((R)0, M_L), supL);
    const R M4R = rmix(M4subR, rmin((R)0, M_R), supR);
    R kp = par[P_KP] / f_a;
    kp = kp * rmax((R)1 - sigma * Mbar2, (R)0);
    kp = kp * (p_R - p_L);
    kp = kp / (rho_i * a_i2);
    const R M_interface = (M4L + M4R) - kp;
    const R alpha4 = (R)0.75 * ((R)-4 + ((R)5 * f_a) * f_a);
    const R P5subL = M2pL * (((R)2 - M_L) + (alpha4 * M_L) * M2mL);
    const R P5subR = M2mR * (((R)-2 - M_R) - (alpha4 * M_R) * M2pR);
    const R P5L = rmix(P5subL, rstep((R)0, M_L), supL);
    const R P5R = rmix(P5subR, rstep((R)0, -M_R), supR);
    R ku = par[P_KU] * P5L;
    ku = ku * P5R;
    ku = ku * (rho_L + rho_R);
    ku = ku * f_a;
    ku = ku * a_i;
    ku = ku * (u_R - u_L);
    P_interface = (P5L * p_L + P5R * p_R) - ku;
    mdot = (a_i * M_interface) * (M_interface > (R)0 ? rho_L : rho_R);
  } else {
    /* ausm.js:348-384 */
    const R a_L = rsqrt_((gamma * rmax(p_L, (R)1e-10)) / rmax(rho_L, (R)1e-10));
    const R a_R = rsqrt_((gamma * rmax(p_R, (R)1e-10)) / rmax(rho_R, (R)1e-10));
    const R a_i = (a_L + a_R) * (R)0.5;
    const R M_L = u_L / a_i;
    const R M_R = u_R / a_i;
    const R Uarc = rsqrt_((R)0.5 * (vvL + vvR));
    const R Marc = rmin((R)1, Uarc / a_i);
    const R X = ((R)1 - Marc) * ((R)1 - Marc);
    const R qL = M_L + (R)1, qR = M_R - (R)1;
    const R sqL = ((R)0.5 * qL) * qL, sqR = ((R)0.5 * qR) * qR;
    const R supL = rstep((R)1, rabs(M_L)), supR = rstep((R)1, rabs(M_R));
    const R fpp = rsat((R)0.5 * rmix(sqL * ((R)2 - (R)1 * M_L), (R)1 + (R)1 * rsign(M_L), supL));
    const R fpm = rsat((R)0.5 * rmix(sqR * ((R)2 - (R)-1 * M_R), (R)1 + (R)-1 * rsign(M_R), supR));
    if (scheme == SCHEME_SLAU2) {
      const R t1 = (fpp - fpm) * (p_L - p_R);
      const R t2 = ((Uarc * ((fpp + fpm) - (R)1)) * (rho_L + rho_R)) * a_i;
      P_interface = (R)0.5 * ((t1 + (p_L + p_R)) + t2);
    } else {
      const R t1 = (fpp - fpm) * (p_L - p_R);
      const R t2 = ((R)1 + ((R)1 - X) * ((fpp + fpm) - (R)1)) * (p_L + p_R);
      P_interface = (R)0.5 * (t1 + t2);
    }
    const R delta_p = p_R - p_L;
    const R ubar_n = (rho_L * rabs(u_L) + rho_R * rabs(u_R)) / (rho_L + rho_R);
    const R g = rsat(-M_L) * rsat(M_R);
    const R ubp = rmix(ubar_n, rabs(u_L), g);
    const R ubm = rmix(ubar_n, rabs(u_R), g);
    mdot = (R)0.5 * ((rho_L * (u_L + ubp) + rho_R * (u_R - ubm)) * ((R)1 - g) - (X / a_i) * delta_p);
  }
  /* convective flux upwinded on the sign of the mass flux, plus the
   * pressure flux (ausm.js:327-335 / 386-391) */
  const int left = mdot > (R)0;
  const R ph1 = left ? vxL : vxR;
  const R ph2 = left ? vyL : vyR;
  const R ph3 = left ? h_L : h_R;
  F[0] = (R)1 * mdot + (R)0;
  F[1] = ph1 * mdot + nx * P_interface;
  F[2] = ph2 * mdot + ny * P_interface;
  F[3] = ph3 * mdot + (R)0;
}

void FN(face_flux)(const R* par, int scheme, int limiter, const R* QL2, const R* QL,
                   const R* QR, const R* QR2, R nx, R ny, R* F) {
  face_flux_frz(par, scheme, limiter, QL2, QL, QR, QR2, nx, ny, F, 0, NULL, NULL);
}

/* face unit normal from two vertices: normalize then rotate +90 degrees */
static inline void face_normal(R dx, R dy, R* nx, R* ny) {
  R len = rsqrt_(dx * dx + dy * dy);
  R tx = dx / len, ty = dy / len;
  *nx = -ty;
  *ny = tx;
}

/* flux pass (ausm.js:178-267 + scheme).  vertical != 0: faces of constant j,
 * texture M x (N+1), face (x,y) below interior cell y.  vertical == 0: faces
 * of constant i, texture (M+1) x N, face (x,y) left of cell x; column M is
 * left zero (the reference computes it but never reads it). */
void FN(flux_pass)(const R* par, int scheme, int limiter, int vertical, int M, int N,
                   const R* vtx, const R* st, R* flux) {
  const int frz = limiter == LIM_NONE ? 0 : FN(frz_mode);
  R* const psi = FN(frz_psi);
  if (vertical) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y <= N; ++y)
      for (int x = 0; x < M; ++x) {
        const R* v1 = vtx + ((size_t)y * M + x) * 2;
        const R* v2 = vtx + ((size_t)y * M + (x + 1) % M) * 2;
        R nx, ny;
        face_normal(v1[0] - v2[0], v1[1] - v2[1], &nx, &ny);
        R* ps = frz ? psi + ((size_t)y * M + x) * 12 : NULL;
        face_flux_frz(par, scheme, limiter, load_state(st, M, x, y - 2), load_state(st, M, x, y - 1),
                      load_state(st, M, x, y), load_state(st, M, x, y + 1), nx, ny,
                      flux + ((size_t)y * M + x) * 4, frz, ps, ps);
      }
  } else {
    const int W = M + 1;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < N; ++y) {
      for (int x = 0; x < M; ++x) {
        const R* v1 = vtx + ((size_t)y * M + x) * 2;
        const R* v2 = vtx + ((size_t)(y + 1) * M + x) * 2;
        R nx, ny;
        face_normal(v2[0] - v1[0], v2[1] - v1[1], &nx, &ny);
        R* ps = frz ? psi + ((size_t)(N + 1) * M + (size_t)y * M + x) * 12 : NULL;
        face_flux_frz(par, scheme, limiter, load_state(st, M, x - 2, y), load_state(st, M, x - 1, y),
                      load_state(st, M, x, y), load_state(st, M, x + 1, y), nx, ny,
                      flux + ((size_t)y * W + x) * 4, frz, ps, ps);
      }
      for (int c = 0; c < 4; ++c) flux[((size_t)y * W + M) * 4 + c] = (R)0;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* boundary pass: ausm.js:34-120                                             */
/* ------------------------------------------------------------------------ */
void FN(boundary_pass)(const R* par, int M, int N, const R* vtx, const R* in, R* out) {
  const R gamma = par[P_GAMMA];
  const R Vx = par[P_VX], Vy = par[P_VY];
#pragma omp parallel for schedule(static)
  for (int y = 0; y < N + 3; ++y)
    for (int x = 0; x < M; ++x) {
      const int rightIdx = (x + 1) % M;
      const int vy = y < N ? y : N;
      const R* v1 = vtx + ((size_t)vy * M + x) * 2;
      const R* v2 = vtx + ((size_t)vy * M + rightIdx) * 2;
      R nx, ny;
      face_normal(v1[0] - v2[0], v1[1] - v2[1], &nx, &ny);
      const R* src = in + ((size_t)y * M + x) * 4;
      R g[4] = {src[0], src[1], src[2], src[3]};
      if (y == 0) {
        const R* I = in + ((size_t)1 * M + x) * 4;
        const R rho = I[0];
        const R ux = I[1] / rho, uy = I[2] / rho;
        const R k = (R)2 * (nx * ux + ny * uy);
        const R gx = ux - k * nx, gy = uy - k * ny;
        g[0] = rho;
        g[1] = gx * rho;
        g[2] = gy * rho;
        g[3] = I[3];
      } else if (y > N) {
        const R* I = in + ((size_t)N * M + x) * 4;
        const R rho = I[0];
        const R ux = I[1] / rho, uy = I[2] / rho;
        const R bio = nx * Vx + ny * Vy;
        if (bio < (R)0) {
          for (int c = 0; c < 4; ++c) g[c] = par[P_IS0 + c];
        } else {
          const R gm1inv = (R)1 / (gamma - (R)1);
          const R u2 = ux * ux + uy * uy;
          const R P = (I[3] - ((R)0.5 * u2) * rho) * (gamma - (R)1);
          const R a2 = (gamma * P) / rho;
          const R a2inf = (gamma * par[P_INP]) / par[P_INRHO];
          const R M_int = u2 / a2;
          const R M_inf2 = (Vx * Vx + Vy * Vy) / a2inf;
          if (M_int < (R)1 && M_inf2 < (R)1) {
            const R un_int = ux * nx + uy * ny;
            const R un_inf = bio;
            const R a_int = rsqrt_(a2);
            const R a_inf = rsqrt_(a2inf);
            const R J_int = un_int + ((R)2 * a_int) * gm1inv;
            const R J_ext = un_inf - ((R)2 * a_inf) * gm1inv;
            const R un_g = (R)0.5 * (J_int + J_ext);
            const R a_g = rmax(a_int * (R)0.1, ((R)0.25 * (J_int - J_ext)) * (gamma - (R)1));
            const R X = rmax((R)1e-6, a_g / a_int);
            const R X2 = X * X;
            const R rho_g = rmax((R)1e-6, ((rho * X2) * X2) * X);
            const R dn = un_g - un_int;
            const R vgx = dn * nx + ux, vgy = dn * ny + uy;
            const R rhoE_g = rmax((R)1e-6, ((rho_g * a_g) * a_g) / gamma) * gm1inv +
                             ((R)0.5 * rho_g) * (vgx * vgx + vgy * vgy);
            g[0] = rho_g;
            g[1] = vgx * rho_g;
            g[2] = vgy * rho_g;
            g[3] = rhoE_g;
          } else if (M_inf2 > (R)1) {
            const R ambientRhoE = par[P_INP] / (gamma - (R)1) + ((R)0.5 * rho) * u2;
            g[0] = I[0];
            g[1] = I[1];
            g[2] = I[2];
            g[3] = ambientRhoE;
          } else {
            for (int c = 0; c < 4; ++c) g[c] = I[c];
          }
          if (y == N + 2)
            for (int c = 0; c < 4; ++c) g[c] = (R)2 * g[c] - I[c];
        }
      }
      R* dst = out + ((size_t)y * M + x) * 4;
      for (int c = 0; c < 4; ++c) dst[c] = g[c];
    }
}

/* ------------------------------------------------------------------------ */
/* residual pass: ausm.js:415-436                                            */
/* ------------------------------------------------------------------------ */
void FN(residual_pass)(int M, int N, const R* fluxX, const R* fluxY, const R* flen,
                       const R* area, R* res) {
  const int W = M + 1;
#pragma omp parallel for schedule(static)
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < M; ++x) {
      const int r = (x + 1) % M;
      const R* FL = fluxX + ((size_t)y * W + x) * 4;
      const R* FR = fluxX + ((size_t)y * W + r) * 4;
      const R* GU = fluxY + ((size_t)(y + 1) * M + x) * 4;
      const R* GD = fluxY + ((size_t)y * M + x) * 4;
      const R* fl = flen + ((size_t)y * M + x) * 4;
      const R A = area[(size_t)y * M + x];
      R* o = res + ((size_t)y * M + x) * 4;
      for (int c = 0; c < 4; ++c) {
        R s = ((FR[c] * fl[1] - FL[c] * fl[0]) + GU[c] * fl[2]) - GD[c] * fl[3];
        o[c] = (-s) / A;
      }
    }
}

/* dt from the max wave speed (ausm.js:457,487,518) */
R FN(dt_from_ws)(const R* par, R maxws) { return rmin(par[P_MAXDT], par[P_CFL] / maxws); }

/* TVD-RK3 stage updates (ausm.js:441-528); interior rows only.
 * k=1: out = s0 + dt*res
 * k=2: out = 0.75*s0 + 0.25*(s1 + dt*res)
 * k=3: out = (s0 + 2*(s1 + dt*res)) / 3                                  */
static void stage_cells(int k, int M, int N, R dt, const R* dtc, const R* res, const R* s0,
                        const R* s1, R* out) {
#pragma omp parallel for schedule(static)
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < M; ++x) {
      const size_t ci = ((size_t)(y + 1) * M + x) * 4;
      const R* rr = res + ((size_t)y * M + x) * 4;
      const R h = dtc ? dtc[(size_t)y * M + x] : dt;
      for (int c = 0; c < 4; ++c) {
        R v;
        if (k == 1) v = s0[ci + c] + h * rr[c];
        else if (k == 2) v = (R)0.75 * s0[ci + c] + (R)0.25 * (s1[ci + c] + h * rr[c]);
        else v = (s0[ci + c] + (R)2 * (s1[ci + c] + h * rr[c])) / (R)3;
        out[ci + c] = v;
      }
    }
}

void FN(stage_pass)(int k, int M, int N, R dt, const R* res, const R* s0, const R* s1, R* out) {
  stage_cells(k, M, N, dt, NULL, res, s0, s1, out);
}

/* wave speeds for CFL (ausm.js:617-630) and their max (ausm.js:555-580) */
R FN(max_wave_speed)(const R* par, int M, int N, const R* st, const R* flen, const R* area,
                     R* ws_out) {
  const R gamma = par[P_GAMMA];
  R mx = (R)0;
#pragma omp parallel for schedule(static) reduction(max : mx)
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < M; ++x) {
      const R* s = st + ((size_t)(y + 1) * M + x) * 4;
      const R rho = s[0];
      const R vx = s[1] / rho, vy = s[2] / rho;
      const R vv = vx * vx + vy * vy;
      const R p = (s[3] - ((R)0.5 * rho) * vv) * (gamma - (R)1);
      const R T = p / rho;
      const R a = rsqrt_(gamma * T);
      const R speed = rsqrt_(vv);
      const R* fl = flen + ((size_t)y * M + x) * 4;
      const R S = ((fl[0] + fl[1]) + fl[2]) + fl[3];
      const R ws = ((speed + a) * S) / ((R)2 * area[(size_t)y * M + x]);
      if (ws_out) ws_out[(size_t)y * M + x] = ws;
      mx = rmax(mx, ws);
    }
  return mx;
}

/* ------------------------------------------------------------------------ */
/* grid: shaders/grid.js                                                     */
/* ------------------------------------------------------------------------ */

/* gridFinalize (shaders/grid.js:128-147): quad area and face lengths */
void FN(grid_finalize)(int M, int N, const R* vtx, R* area, R* flen) {
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < M; ++x) {
      const int r = (x + 1) % M;
      const R* a = vtx + ((size_t)y * M + x) * 2;
      const R* b = vtx + ((size_t)y * M + r) * 2;
      const R* c = vtx + ((size_t)(y + 1) * M + x) * 2;
      const R* d = vtx + ((size_t)(y + 1) * M + r) * 2;
      const R cr = (d[0] - a[0]) * (c[1] - b[1]) - (c[0] - b[0]) * (d[1] - a[1]);
      area[(size_t)y * M + x] = (R)0.5 * rabs(cr);
      R* fl = flen + ((size_t)y * M + x) * 4;
      R ex, ey;
      ex = a[0] - c[0]; ey = a[1] - c[1]; fl[0] = rsqrt_(ex * ex + ey * ey);
      ex = b[0] - d[0]; ey = b[1] - d[1]; fl[1] = rsqrt_(ex * ex + ey * ey);
      ex = c[0] - d[0]; ey = c[1] - d[1]; fl[2] = rsqrt_(ex * ex + ey * ey);
      ex = a[0] - b[0]; ey = a[1] - b[1]; fl[3] = rsqrt_(ex * ex + ey * ey);
    }
}

/* gridInterpolation (shaders/grid.js:15-28): linear blend body -> far field
 * along each i line.  vtx rows 0 and N must hold the two boundaries. */
void FN(grid_interpolate)(int M, int N, R* vtx) {
  for (int y = 1; y < N; ++y) {
    const R t = (R)y / (R)N;
    for (int x = 0; x < M; ++x) {
      const R* o = vtx + (size_t)x * 2;
      const R* b = vtx + ((size_t)N * M + x) * 2;
      R* p = vtx + ((size_t)y * M + x) * 2;
      p[0] = rmix(o[0], b[0], t);
      p[1] = rmix(o[1], b[1], t);
    }
  }
}

/* One Jacobi sweep of the Thompson-Thames-Mastin / Winslow equations with
 * P = Q = 0 (shaders/grid.js:44-110).  Rows 0 and N are fixed; i wraps.
 * A point moves only if the local Jacobian is negative (valid orientation). */
void FN(grid_jacobi_sweep)(int M, int N, const R* in, R* out) {
  memcpy(out, in, sizeof(R) * 2 * (size_t)M);
  memcpy(out + (size_t)N * M * 2, in + (size_t)N * M * 2, sizeof(R) * 2 * (size_t)M);
  for (int y = 1; y < N; ++y)
    for (int x = 0; x < M; ++x) {
      const int l = (x + M - 1) % M, r = (x + 1) % M;
#define V(xx, yy, c) in[((size_t)(yy) * M + (xx)) * 2 + (c)]
      R xi[2], eta[2], xe[2], ctr[2], lr[2], du[2];
      for (int c = 0; c < 2; ++c) {
        ctr[c] = V(x, y, c);
        xi[c] = (R)0.5 * (V(r, y, c) - V(l, y, c));
        eta[c] = (R)0.5 * (V(x, y + 1, c) - V(x, y - 1, c));
        xe[c] = (R)0.25 * (((V(r, y + 1, c) + V(l, y - 1, c)) - V(l, y + 1, c)) - V(r, y - 1, c));
        lr[c] = V(l, y, c) + V(r, y, c);
        du[c] = V(x, y - 1, c) + V(x, y + 1, c);
      }
#undef V
      const R alpha = eta[0] * eta[0] + eta[1] * eta[1];
      const R gam = xi[0] * xi[0] + xi[1] * xi[1];
      const R beta = xi[0] * eta[0] + xi[1] * eta[1];
      const R jac = xi[0] * eta[1] - xi[1] * eta[0];
      R* o = out + ((size_t)y * M + x) * 2;
      for (int c = 0; c < 2; ++c) {
        /* P = Q = 0: the source terms add P*xy_xi = 0 (shaders/grid.js:102) */
        const R np = (alpha * (lr[c] + (R)0 * xi[c]) + gam * (du[c] + (R)0 * eta[c]) -
                      ((R)2 * beta) * xe[c]) / ((R)2 * (alpha + gam));
        o[c] = jac < (R)0 ? rmix(ctr[c], np, (R)1) : ctr[c];
      }
    }
}

/* Jacobi iterations with ping-pong; result in vtx.  Returns nothing. */
void FN(grid_jacobi)(int M, int N, R* vtx, int iters) {
  R* tmp = (R*)malloc(sizeof(R) * 2 * (size_t)M * (N + 1));
  for (int it = 0; it < iters; ++it) {
    FN(grid_jacobi_sweep)(M, N, vtx, tmp);
    memcpy(vtx, tmp, sizeof(R) * 2 * (size_t)M * (N + 1));
  }
  free(tmp);
}

/* ------------------------------------------------------------------------ */
/* forces: shaders/graphs.js:16-101                                          */
/* ------------------------------------------------------------------------ */

/* Per-segment surface force p*len*(-n_out) from row-1 pressure
 * (graphs.js:20-39), summed in index order in FP64; out[0..1] = force,
 * out[2] = Cl, out[3] = Cd, both divided by q and ref_len (the reference
 * does not divide by chord: ref_len = 1 reproduces it).  cp[M] (optional)
 * gets (p - p_inf)/q per segment. */
void FN(forces)(const R* par, int M, const R* st, const R* vtx, double ref_len, double* out,
                double* cp) {
  const R gamma = par[P_GAMMA];
  double fx = 0, fy = 0;
  const double Vx = par[P_VX], Vy = par[P_VY];
  const double q = 0.5 * (double)par[P_INRHO] * (Vx * Vx + Vy * Vy);
  for (int i = 0; i < M; ++i) {
    const R* s = st + ((size_t)1 * M + i) * 4;
    const R rho = s[0];
    const R vx = s[1] / rho, vy = s[2] / rho;
    const R p = (s[3] - ((R)0.5 * rho) * (vx * vx + vy * vy)) * (gamma - (R)1);
    const int r = (i + 1) % M;
    const R* v1 = vtx + (size_t)i * 2;
    const R* v2 = vtx + (size_t)r * 2;
    const R ex = v2[0] - v1[0], ey = v2[1] - v1[1];
    const R fl = rsqrt_(ex * ex + ey * ey);
    const R nx = (v2[1] - v1[1]) / fl, ny = (v1[0] - v2[0]) / fl;
    const R pf = p * fl;
    fx += (double)(pf * -nx);
    fy += (double)(pf * -ny);
    if (cp) cp[i] = ((double)p - (double)par[P_INP]) / q;
  }
  const double vn = sqrt(Vx * Vx + Vy * Vy);
  const double ux = Vx / vn, uy = Vy / vn;
  out[0] = fx;
  out[1] = fy;
  out[2] = (fx * -uy + fy * ux) / (q * ref_len);
  out[3] = (fx * ux + fy * uy) / (q * ref_len);
}

/* ------------------------------------------------------------------------ */
/* step driver: main.js:744-792                                              */
/* ------------------------------------------------------------------------ */

/* Fills every texel with the uniform state (prepareState, ausm.js:1-18, in
 * "rest" mode; the freestream mode writes inState instead). */
void FN(init_state)(const R* par, int M, int N, int freestream, R* st) {
  const R rest[4] = {par[P_INRHO], (R)0, (R)0, par[P_INP] / (par[P_GAMMA] - (R)1)};
  const R* q = freestream ? par + P_IS0 : rest;
  for (size_t i = 0; i < (size_t)M * (N + 3); ++i)
    for (int c = 0; c < 4; ++c) st[i * 4 + c] = q[c];
}

/* Residual L(Q) of one state (boundary + both flux passes + residual), the
 * quantity each RK stage adds.  st_in: (N+3)xMx4, res: NxMx4. */
void FN(eval_residual)(const R* par, int scheme, int limiter, int M, int N, const R* vtx,
                       const R* flen, const R* area, const R* st_in, R* res) {
  R* sb = (R*)malloc(sizeof(R) * 4 * (size_t)M * (N + 3));
  R* fx = (R*)malloc(sizeof(R) * 4 * (size_t)(M + 1) * N);
  R* fy = (R*)malloc(sizeof(R) * 4 * (size_t)M * (N + 1));
  FN(boundary_pass)(par, M, N, vtx, st_in, sb);
  FN(flux_pass)(par, scheme, limiter, 1, M, N, vtx, sb, fy);
  FN(flux_pass)(par, scheme, limiter, 0, M, N, vtx, sb, fx);
  FN(residual_pass)(M, N, fx, fy, flen, area, res);
  free(sb);
  free(fx);
  free(fy);
}

/* Runs nsteps TVD-RK3 steps in the reference's 15-pass structure.
 * s0, s1, s2 are the three state textures; the current solution is in s2 on
 * entry and on exit (as after prepareState, main.js:631-655).  *maxws holds
 * the max wave speed the next step uses; it is recomputed from s2 after
 * every `refresh` steps (the reference refreshes once per frame of
 * stepsPerFrame steps, main.js:789-791).  step_count carries the position in
 * that cadence across calls.  res (optional, NxMx4) receives the stage-3
 * residual of the last step.  Returns the dt of the last step.
 *
 * local_dt != 0 (local time stepping, a build extension with no reference
 * counterpart): every cell steps with its own dt = dt_from_ws(wave speed of
 * the cell in Q^n), the same in all three stages; *maxws and the returned
 * global dt are kept as in the global mode. */
R FN(run_steps_dt)(const R* par, int scheme, int limiter, int M, int N, const R* vtx,
                   const R* flen, const R* area, R* s0, R* s1, R* s2, int nsteps, int refresh,
                   long long* step_count, R* maxws, R* res_out, int local_dt) {
  R* fx = (R*)malloc(sizeof(R) * 4 * (size_t)(M + 1) * N);
  R* fy = (R*)malloc(sizeof(R) * 4 * (size_t)M * (N + 1));
  R* res = (R*)malloc(sizeof(R) * 4 * (size_t)M * N);
  R* dtc = local_dt ? (R*)malloc(sizeof(R) * (size_t)M * N) : NULL;
  R dt = 0;
  for (int s = 0; s < nsteps; ++s) {
    dt = FN(dt_from_ws)(par, *maxws);
    if (dtc) {
      (void)FN(max_wave_speed)(par, M, N, s2, flen, area, dtc);
      for (size_t c = 0; c < (size_t)M * N; ++c) dtc[c] = FN(dt_from_ws)(par, dtc[c]);
    }
    /* stage 1: state2 -> state0 (Qn), Q1 -> state1 */
    FN(boundary_pass)(par, M, N, vtx, s2, s0);
    FN(flux_pass)(par, scheme, limiter, 1, M, N, vtx, s0, fy);
    FN(flux_pass)(par, scheme, limiter, 0, M, N, vtx, s0, fx);
    FN(residual_pass)(M, N, fx, fy, flen, area, res);
    stage_cells(1, M, N, dt, dtc, res, s0, NULL, s1);
    /* stage 2: state1 -> state2 (Q1), Q2 -> state2 */
    FN(boundary_pass)(par, M, N, vtx, s1, s2);
    FN(flux_pass)(par, scheme, limiter, 1, M, N, vtx, s2, fy);
    FN(flux_pass)(par, scheme, limiter, 0, M, N, vtx, s2, fx);
    FN(residual_pass)(M, N, fx, fy, flen, area, res);
    stage_cells(2, M, N, dt, dtc, res, s0, s1, s2);
    /* stage 3: state2 -> state1 (Q2), Qn+1 -> state2 */
    FN(boundary_pass)(par, M, N, vtx, s2, s1);
    FN(flux_pass)(par, scheme, limiter, 1, M, N, vtx, s1, fy);
    FN(flux_pass)(par, scheme, limiter, 0, M, N, vtx, s1, fx);
    FN(residual_pass)(M, N, fx, fy, flen, area, res);
    stage_cells(3, M, N, dt, dtc, res, s0, s1, s2);
    *step_count += 1;
    if (refresh > 0 && (*step_count % refresh) == 0)
      *maxws = FN(max_wave_speed)(par, M, N, s2, flen, area, NULL);
  }
  if (res_out) memcpy(res_out, res, sizeof(R) * 4 * (size_t)M * N);
  free(fx);
  free(fy);
  free(res);
  free(dtc);
  return dt;
}

R FN(run_steps)(const R* par, int scheme, int limiter, int M, int N, const R* vtx,
                const R* flen, const R* area, R* s0, R* s1, R* s2, int nsteps, int refresh,
                long long* step_count, R* maxws, R* res_out) {
  return FN(run_steps_dt)(par, scheme, limiter, M, N, vtx, flen, area, s0, s1, s2, nsteps, refresh,
                          step_count, maxws, res_out, 0);
}

/* Compensated-state TVD-RK3 (a build extension with no reference
 * counterpart).  The same Shu-Osher scheme in increment form,
 *   K_s = dt L(Q_{s-1}),  Q1 = Q^n + K1,  Q2 = Q^n + (K1 + K2)/4,
 *   Q^{n+1} = Q^n + (K1 + K2 + 4 K3)/6,
 * with the solution held as the unevaluated sum hi + lo of two R values:
 * the final increment is added with an error-free TwoSum and the rounding
 * error carried in lo, so that increments below half an ulp of hi are not
 * rounded away (the stagnation of plain R accumulation near a steady
 * state).  Fluxes see hi only (round(hi + lo) = hi).  s0/s1/s2 are the
 * three state textures as in run_steps (s2 holds hi on entry and exit);
 * lo is (N+3) x M x 4 (interior rows used) and acc N x M x 4 scratch.
 * Per cell and component, in this order:
 *   stage 1: K = dt*res;  acc = K;                 Q1 = hi + K
 *   stage 2: K = dt*res;  acc = acc + K;           Q2 = hi + 0.25*acc
 *   stage 3: K = dt*res;  D = (acc + 4*K) / 6;
 *            s = hi + D;  b = s - hi;  e = (hi - (s - b)) + (D - b);
 *            t = lo + e;  h = s + t;   b2 = h - s;
 *            lo = (s - (h - b2)) + (t - b2);  hi = h. */
static void comp_stage(int k, int M, int N, R dt, const R* dtc, const R* res, const R* hi, R* lo,
                       R* acc, R* out) {
#pragma omp parallel for schedule(static)
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < M; ++x) {
      const size_t ci = ((size_t)(y + 1) * M + x) * 4;
      const size_t ri = ((size_t)y * M + x) * 4;
      const R h = dtc ? dtc[(size_t)y * M + x] : dt;
      for (int c = 0; c < 4; ++c) {
        const R K = h * res[ri + c];
        if (k == 1) {
          acc[ri + c] = K;
          out[ci + c] = hi[ci + c] + K;
        } else if (k == 2) {
          const R a = acc[ri + c] + K;
          acc[ri + c] = a;
          out[ci + c] = hi[ci + c] + (R)0.25 * a;
        } else {
          const R D = (acc[ri + c] + (R)4 * K) / (R)6;
          const R q = hi[ci + c];
          const R sm = q + D;
          const R b = sm - q;
          const R e = (q - (sm - b)) + (D - b);
          const R t = lo[ci + c] + e;
          const R hn = sm + t;
          const R b2 = hn - sm;
          lo[ci + c] = (sm - (hn - b2)) + (t - b2);
          out[ci + c] = hn;
        }
      }
    }
}

R FN(run_steps_comp)(const R* par, int scheme, int limiter, int M, int N, const R* vtx,
                     const R* flen, const R* area, R* s0, R* s1, R* s2, R* lo, int nsteps,
                     int refresh, long long* step_count, R* maxws, int local_dt) {
  R* fx = (R*)malloc(sizeof(R) * 4 * (size_t)(M + 1) * N);
  R* fy = (R*)malloc(sizeof(R) * 4 * (size_t)M * (N + 1));
  R* res = (R*)malloc(sizeof(R) * 4 * (size_t)M * N);
  R* acc = (R*)malloc(sizeof(R) * 4 * (size_t)M * N);
  R* hi = (R*)malloc(sizeof(R) * 4 * (size_t)M * (N + 3));
  R* dtc = local_dt ? (R*)malloc(sizeof(R) * (size_t)M * N) : NULL;
  R dt = 0;
  for (int s = 0; s < nsteps; ++s) {
    dt = FN(dt_from_ws)(par, *maxws);
    if (dtc) {
      (void)FN(max_wave_speed)(par, M, N, s2, flen, area, dtc);
      for (size_t c = 0; c < (size_t)M * N; ++c) dtc[c] = FN(dt_from_ws)(par, dtc[c]);
    }
    memcpy(hi, s2, sizeof(R) * 4 * (size_t)M * (N + 3));
    /* stage 1: hi -> s0 (with ghosts), Q1 -> s1 */
    FN(boundary_pass)(par, M, N, vtx, hi, s0);
    FN(flux_pass)(par, scheme, limiter, 1, M, N, vtx, s0, fy);
    FN(flux_pass)(par, scheme, limiter, 0, M, N, vtx, s0, fx);
    FN(residual_pass)(M, N, fx, fy, flen, area, res);
    comp_stage(1, M, N, dt, dtc, res, hi, lo, acc, s1);
    /* stage 2: Q1 -> Q2 */
    FN(boundary_pass)(par, M, N, vtx, s1, s2);
    FN(flux_pass)(par, scheme, limiter, 1, M, N, vtx, s2, fy);
    FN(flux_pass)(par, scheme, limiter, 0, M, N, vtx, s2, fx);
    FN(residual_pass)(M, N, fx, fy, flen, area, res);
    comp_stage(2, M, N, dt, dtc, res, hi, lo, acc, s2);
    /* stage 3: Q2 -> (hi, lo) */
    FN(boundary_pass)(par, M, N, vtx, s2, s1);
    FN(flux_pass)(par, scheme, limiter, 1, M, N, vtx, s1, fy);
    FN(flux_pass)(par, scheme, limiter, 0, M, N, vtx, s1, fx);
    FN(residual_pass)(M, N, fx, fy, flen, area, res);
    comp_stage(3, M, N, dt, dtc, res, hi, lo, acc, s2);
    *step_count += 1;
    if (refresh > 0 && (*step_count % refresh) == 0)
      *maxws = FN(max_wave_speed)(par, M, N, s2, flen, area, NULL);
  }
  free(fx);
  free(fy);
  free(res);
  free(acc);
  free(hi);
  free(dtc);
  return dt;
}

/* Derived display fields of the visualization pass (ausm.js:586-710), the
 * scalar before the colour map, for every interior cell (out N x M), and the
 * free-stream value the colour map centres on (*fs).  mode = simDisplayMode:
 * 0 schlieren exp(-2|grad rho|) with the pass's one-sided clamps (down row
 * max(j-1, 0), up row min(j+1, N-2)), 1 vorticity dv/dx - du/dy by the same
 * differences, 2 density, 3 temperature p/rho, 4 pressure, 5 and 6 Mach
 * (free-stream value 0.5 or the inflow Mach), 7 |v - V_inf| (the pass shows
 * its components), 8 entropy log(p / rho^gamma), 9 total pressure.
 * Parity: modes 1-7 bitwise (division and square root only), 0, 8, 9 to a
 * few ulp of exp / log / pow. */
void FN(derived_field)(const R* par, int M, int N, const R* st, int mode, R* out, R* fs) {
  const R gamma = par[P_GAMMA];
  const R inRho = par[P_INRHO], inP = par[P_INP];
  const R Vx = par[P_VX], Vy = par[P_VY];
  const R inVel = rsqrt_(Vx * Vx + Vy * Vy);
  const R inTemp = inP / inRho;
  const R inMach = inVel / rsqrt_(gamma * inTemp);
  const int fl = sizeof(R) == 4;
  switch (mode) {
    case 0: *fs = (R)1; break;
    case 1: case 7: *fs = (R)0; break;
    case 3: *fs = inTemp; break;
    case 4: *fs = inP; break;
    case 5: *fs = (R)0.5; break;
    case 6: *fs = inMach; break;
    case 8: *fs = fl ? (R)logf((float)(inP / (R)powf((float)inRho, (float)gamma)))
                     : (R)log((double)(inP / (R)pow((double)inRho, (double)gamma))); break;
    case 9: {
      const R b = (R)1 + (R)0.5 * (gamma - (R)1) * inMach * inMach;
      const R e = gamma / (gamma - (R)1);
      *fs = inP * (fl ? (R)powf((float)b, (float)e) : (R)pow((double)b, (double)e));
      break;
    }
    default: *fs = inRho;
  }
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < M; ++x) {
      const R* s = st + ((size_t)(y + 1) * M + x) * 4;
      const R rho = s[0];
      const R vx = s[1] / rho, vy = s[2] / rho;
      const R vv = vx * vx + vy * vy;
      const R p = (s[3] - ((R)0.5 * rho) * vv) * (gamma - (R)1);
      const R T = p / rho;
      const R a = rsqrt_(gamma * T);
      const R speed = rsqrt_(vv);
      R v;
      if (mode <= 1) {
        const R* L = st + ((size_t)(y + 1) * M + (x + M - 1) % M) * 4;
        const R* Rr = st + ((size_t)(y + 1) * M + (x + 1) % M) * 4;
        const int dn = y > 1 ? y : 1;                 /* texture rows */
        const int upr = y + 2 < N - 1 ? y + 2 : N - 1;
        const R* D = st + ((size_t)dn * M + x) * 4;
        const R* U = st + ((size_t)upr * M + x) * 4;
        if (mode == 0) {
          const R gx = Rr[0] - L[0], gy = U[0] - D[0];
          const R g = rsqrt_(gx * gx + gy * gy);
          v = fl ? (R)expf((float)(-g * (R)2)) : (R)exp((double)(-g * (R)2));
        } else {
          v = (Rr[2] / Rr[0] - L[2] / L[0]) - (U[1] / U[0] - D[1] / D[0]);
        }
      } else {
        switch (mode) {
          case 3: v = T; break;
          case 4: v = p; break;
          case 5: case 6: v = speed / a; break;
          case 7: {
            const R dx = vx - Vx, dy = vy - Vy;
            v = rsqrt_(dx * dx + dy * dy);
            break;
          }
          case 8:
            v = fl ? (R)logf((float)(p / (R)powf((float)rho, (float)gamma)))
                   : (R)log((double)(p / (R)pow((double)rho, (double)gamma)));
            break;
          case 9: {
            const R m = speed / a;
            const R b = (R)1 + (R)0.5 * (gamma - (R)1) * m * m;
            const R e = gamma / (gamma - (R)1);
            v = p * (fl ? (R)powf((float)b, (float)e) : (R)pow((double)b, (double)e));
            break;
          }
          default: v = rho;
        }
      }
      out[(size_t)y * M + x] = v;
    }
}

/* L2 norm of one residual component over the interior, accumulated in FP64
 * in index order. */
double FN(residual_norm)(int M, int N, const R* res, int comp) {
  double s = 0;
  for (size_t i = 0; i < (size_t)M * N; ++i) {
    const double v = (double)res[i * 4 + comp];
    s += v * v;
  }
  return sqrt(s);
}
