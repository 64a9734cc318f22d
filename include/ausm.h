/*
 * ausm.h — C ABI of the MI355X-native explicit O-grid solver step.
 *
 * Drop-in replacement for the reference's WebGPU compute passes
 * (huj31415/ausm-webgpu): grid + free stream + scheme/limiter/CFL in,
 * conserved-state buffer out.  Plain C types only; every call returns an
 * int status (AUSM_OK = 0, negative on error) and ausm_last_error() gives a
 * thread-local message.  Host arrays passed in are copied, never retained.
 *
 * Array layouts are the reference's texture layouts, row-major with the i
 * index (around the body) fastest:
 *   vertices  (N+1) x M x 2   gridPoints  (reference src/main.js:110)
 *   state     (N+3) x M x 4   state0/1/2  (src/main.js:117-119): row 0 is the
 *             body ghost, rows 1..N the interior cells, rows N+1, N+2 the
 *             far-field ghosts; (rho, rho*u, rho*v, rho*E)
 *   residual  N x M x 4       (src/main.js:123)
 *   area      N x M           (src/main.js:113)
 *   face len  N x M x 4       (left, right, up, down)  (src/main.js:114)
 *
 * Threading: a context belongs to one host thread at a time; separate
 * contexts may be used concurrently.  Each context owns one HIP stream and
 * all of its device buffers.
 */
#ifndef AUSM_H
#define AUSM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AUSM_VERSION 1

/* status codes */
#define AUSM_OK 0
#define AUSM_ERR_ARG (-1)       /* bad argument */
#define AUSM_ERR_HIP (-2)       /* HIP runtime error */
#define AUSM_ERR_NOMEM (-3)     /* device or host allocation failed */
#define AUSM_ERR_ORDER (-4)     /* call out of order (no grid yet, ...) */
#define AUSM_ERR_NONFINITE (-5) /* NaN/Inf detected in the solution */
#define AUSM_ERR_FULL (-6)      /* batch: no free slot or memory */
#define AUSM_ERR_NOTFOUND (-7)  /* batch: unknown solve id */

/* flux schemes, in the order of the reference's `solvers` table
 * (src/util.js:114-118) */
enum ausm_scheme { AUSM_SLAU2 = 0, AUSM_SLAU = 1, AUSM_AUSMUP = 2 };

/* MUSCL limiters.  VANALBADA is the one the reference wires
 * (src/shaders/ausm.js:247-248); MINMOD and VANLEER are the reference's
 * unwired limiter functions (ausm.js:159-166); NONE is uni.muscl = 0
 * (first order, ausm.js:240).  VANALBADA_CLAMPED has no counterpart in the
 * reference: van Albada as a TVD limiter, phi(r) = (r^2+r)/(r^2+1) for
 * r > 0 and 0 for r <= 0 (the reference's form goes negative for
 * -1 < r < 0), a convergence option for steady solves. */
enum ausm_limiter {
  AUSM_VANALBADA = 0, AUSM_MINMOD = 1, AUSM_VANLEER = 2, AUSM_NOLIMITER = 3,
  AUSM_VANALBADA_CLAMPED = 4
};

/* arithmetic of the stage kernels: EXACT follows the reference WGSL
 * expression by expression with IEEE division/sqrt and no contraction;
 * FAST shares reciprocals, uses hardware rcp/sqrt and FMA contraction. */
enum ausm_math { AUSM_MATH_EXACT = 0, AUSM_MATH_FAST = 1 };

/* initial state: REST is the reference's prepareState (ausm.js:1-18, air at
 * rest); FREESTREAM fills every cell with inState (main.js:783-784). */
enum ausm_init { AUSM_INIT_REST = 0, AUSM_INIT_FREESTREAM = 1 };

/* elliptic grid smoother: JACOBI is the reference's ping-pong sweep
 * (src/shaders/grid.js:44-110); SOR is red-black over-relaxation of the same
 * update, each half-sweep relaxing one colour from the previous half-sweep's
 * values (deterministic; needs even M for the colouring to hold across the
 * seam). */
enum ausm_grid_method { AUSM_GRID_JACOBI = 0, AUSM_GRID_SOR = 1 };

/* body-point spacing along the chord */
enum ausm_spacing { AUSM_SPACING_POWER = 0, AUSM_SPACING_COSINE = 1 };

typedef struct ausm_ctx ausm_ctx;

typedef struct ausm_step_info {
  float dt;             /* time step of the last step taken */
  float max_wave_speed; /* max wave speed that dt is currently based on */
  double res_l2;        /* FP64 L2 norm of the density residual L(Q^n) of the
                           last step (stage 1), over all interior cells */
  int nonfinite;        /* 1 if the max wave speed is not finite */
  long long steps_done; /* steps taken on this context so far */
} ausm_step_info;

/* ---- errors / version ---------------------------------------------------- */
const char* ausm_last_error(void);
int ausm_version(void);

/* ---- context (replaces requestDevice + the storage table, main.js:5-151,
 *      util.js:41-74) ---------------------------------------------------- */
int ausm_ctx_create(int device, int M, int N, ausm_ctx** out);
int ausm_ctx_destroy(ausm_ctx* ctx);
/* the context's HIP stream (hipStream_t), for callers that time it */
int ausm_ctx_stream(ausm_ctx* ctx, void** stream);
int ausm_sync(ausm_ctx* ctx);

/* ---- body / far-field generators (src/grid.js:38-52, 232-255), FP64 ----- */
/* NACA 4-digit section with maximum camber m, its position p and thickness
 * t (NACA 2412: m=0.02, p=0.4, t=0.12), closed trailing edge; chord 2*size
 * centred at x_offset; M points CCW from the trailing edge. */
int ausm_boundary_naca4(double m, double p, double t, double size, double x_offset,
                        double exponent, int spacing, int M, double* xy);
/* Class-shape transformation (Kulfan CST) section; no counterpart in the
 * reference, which generates NACA 4-digit bodies only.  n_coef Bernstein
 * coefficients per side (a_lower signed: negative below the chord), class
 * function sqrt(x)(1-x) (round nose of radius a_upper[0]^2/2), trailing-edge
 * gap te_gap (0 = closed); chord stations, point order and placement as
 * ausm_boundary_naca4. */
int ausm_boundary_cst(const double* a_upper, const double* a_lower, int n_coef, double te_gap,
                      double size, double x_offset, double exponent, int spacing, int M,
                      double* xy);
/* Bezier section; no counterpart in the reference.  upper and lower hold
 * n_upper and n_lower (2..32) control points (x, y) in chord units, each
 * surface from the leading edge (t = 0, one point shared by both) to its
 * trailing-edge point (t = 1); the curve parameter takes the chordwise
 * stations of ausm_boundary_naca4 (spacing, exponent), and the chord 2*size
 * is placed as there.  Non-finite control points, or surfaces that do not
 * share the leading-edge point, are AUSM_ERR_ARG. */
int ausm_boundary_bezier(const double* upper, int n_upper, const double* lower, int n_lower,
                         double size, double x_offset, double exponent, int spacing, int M,
                         double* xy);
/* Bodies the reference's object menu offers (src/grid.js:258-263), in the
 * point order of ausm_boundary_naca4:
 *   Sears-Haack body of volume V and length L (generateSearsHaackBoundary,
 *   src/grid.js:53-61); Haack-series projectile: nose fineness F, radius R,
 *   body length L (fraction of the nose length), boat-tail radius and
 *   length fractions Rb, Lb, Haack parameter C (generateHaackProjectileBoundary,
 *   src/grid.js:75-118); regular N-gon of inradius R rotated by a and
 *   squashed by F in y, the reference's circle being N = 400, F = 1
 *   (generatePolygonBoundary, src/grid.js:224-231). */
int ausm_boundary_sears_haack(double V, double L, double x_offset, int M, double* xy);
int ausm_boundary_haack_projectile(double F, double R, double L, double Rb, double Lb, double C,
                                   double x_offset, int M, double* xy);
int ausm_boundary_polygon(double R, int N, double F, double a, double x_offset, int M, double* xy);
/* Airfoil coordinate file (Selig or Lednicer text, auto-detected) sampled
 * at M points with a centripetal Catmull-Rom spline over arc length, denser
 * at the nose; chord 2*size at x_offset (parseDat + sampleAirfoil,
 * src/airfoilDatParser.js:67-293).  format (optional) gets 0 = Selig,
 * 1 = Lednicer. */
int ausm_boundary_dat(const char* text, double size, double x_offset, int M, double* xy,
                      int* format);
int ausm_boundary_circle(double radius, int M, double* xy);
int ausm_boundary_rect(double length_ratio, int M, double* xy);

/* ---- grid (prepareGrid, main.js:658-707) -------------------------------- */
/* interpolation + `iters` elliptic sweeps + finalize, in FP64 on the GPU;
 * omega is used by SOR only. */
int ausm_grid_from_boundary(ausm_ctx* ctx, const double* body_xy, const double* outer_xy,
                            int iters, int method, double omega);
int ausm_grid_set_vertices(ausm_ctx* ctx, const float* xy);
int ausm_grid_get_vertices(ausm_ctx* ctx, float* xy);
int ausm_grid_get_vertices64(ausm_ctx* ctx, double* xy);
/* FP64 vertices (the FP64 solve uses them as given, the FP32 solve rounded;
 * ausm_grid_set_vertices widens its FP32 vertices for the FP64 solve) */
int ausm_grid_set_vertices64(ausm_ctx* ctx, const double* xy);
int ausm_grid_get_metrics(ausm_ctx* ctx, float* area, float* face_lengths);

/* ---- parameters (Uniforms fields, src/util.js:2-27) --------------------- */
/* free stream; inState = (rho, rho*vx, rho*vy, p/(gamma-1) + rho*|v|^2/2) is
 * formed in double and stored as FP32, as main.js:780-784 does. */
int ausm_set_freestream(ausm_ctx* ctx, double rho, double p, double vx, double vy,
                        double gamma);
int ausm_set_scheme(ausm_ctx* ctx, int scheme, int limiter, float K_p, float K_u);
/* dt = min(maxdt, cfl / max wave speed); the wave speed is refreshed from
 * the solution after every `steps_per_dt_refresh` steps counted from the
 * last refresh (the reference refreshes once per frame of stepsPerFrame
 * steps, main.js:789-791; a host that varies stepsPerFrame sets it before
 * each frame). */
int ausm_set_cfl(ausm_ctx* ctx, float cfl, float maxdt, int steps_per_dt_refresh);
int ausm_set_math(ausm_ctx* ctx, int math);
/* Time step.  GLOBAL is the reference's: one dt = min(maxdt, cfl / max
 * wave speed) for every cell (main.js:789-791, ausm.js:457).  LOCAL has no
 * counterpart in the reference: each cell steps with its own
 * min(maxdt, cfl / wave speed) from Q^n, the same in all three RK stages --
 * the usual convergence accelerator for steady solves (the spatial operator
 * and hence the steady state are unchanged; the transient is not
 * time-accurate).  ausm_step_info.dt still reports the global dt. */
enum ausm_dt_mode { AUSM_DT_GLOBAL = 0, AUSM_DT_LOCAL = 1 };
int ausm_set_dt_mode(ausm_ctx* ctx, int mode);
/* rows per wavefront tile of the stage kernel (0 = autotuned) */
int ausm_set_tile(ausm_ctx* ctx, int rows_per_wave);

/* Limiter freezing.  No counterpart in the reference (a convergence device
 * for steady solves with a MUSCL reconstruction, whose limiter switching
 * can hold the residual in a limit cycle).  With r the reference's limiter
 * ratio, psi = phi(r)/r and w = dF^2 / (dF^2 + 1e-6), the limited face
 * values are exactly W_L + psi_L (w dL + (1 - w) dF)/2 and
 * W_R - psi_R (w dR + (1 - w) dF)/2.  on = 1 records psi (van Albada:
 * AUSM_VANALBADA, or AUSM_VANALBADA_CLAMPED with psi = 0 for r <= 0) and w
 * of every face and primitive component from the current state; every
 * later stage, residual and norm reconstructs with the record: the same
 * operator at the recorded state, with the limiter no longer switching.
 * on = 0, a new state, grid, limiter or precision ends it.  Needs a state;
 * other limiters: AUSM_ERR_ARG.  Device memory: 24 planes of
 * (N+1) x round_up(M, 64) doubles, allocated on first use. */
int ausm_freeze_limiter(ausm_ctx* ctx, int on);
int ausm_limiter_frozen(ausm_ctx* ctx, int* on);

/* Arithmetic precision of the solve.  No counterpart in the reference,
 * whose storage and arithmetic are all f32 (FP32 here).  FP64 runs the same
 * scheme and pass structure in double (exact arithmetic, FP64 state, the
 * FP64 grid, the free stream unrounded): the comparison solve that bounds
 * the error of FP32 results.  FP32C keeps every flux and stage operation in
 * FP32 but holds the solution as an unevaluated pair hi + lo and adds each
 * step's increment with an error-free TwoSum (the same TVD-RK3 in increment
 * form, Q^{n+1} = Q^n + (K1 + K2 + 4 K3)/6): plain FP32 accumulation
 * stagnates near a steady state once dt*L(Q) falls below half an ulp of Q,
 * FP32C does not.  Switching carries the current solution over.  In FP64
 * the FP32 accessors (ausm_get_state, ausm_eval_residual, ausm_wave_speeds,
 * ausm_step_info's dt and wave speed) return values rounded to FP32; in
 * FP32C they return hi, and ausm_get_state64 returns hi + lo.  Batch
 * contexts are FP32 only. */
enum ausm_precision { AUSM_PREC_FP32 = 0, AUSM_PREC_FP64 = 1, AUSM_PREC_FP32C = 2 };
int ausm_set_precision(ausm_ctx* ctx, int precision);

/* ---- state and stepping (prepareState / render step loop,
 *      main.js:631-655, 742-792) ------------------------------------------ */
int ausm_init_state(ausm_ctx* ctx, int mode);
int ausm_set_state(ausm_ctx* ctx, const float* q);
/* interior rows as stored; ghost rows as the next boundary pass makes them */
int ausm_get_state(ausm_ctx* ctx, float* q);
int ausm_step(ausm_ctx* ctx, int nsteps, ausm_step_info* info);
/* Captures, instantiates and uploads the step graphs that a following
 * ausm_step(ctx, nsteps, ...) will launch -- from the current point, and
 * from the refresh boundary any earlier step call leaves -- without running
 * them; the graphs stay cached on the context until a setting, the grid or
 * the state changes.  No counterpart in the reference (its command encoder records a
 * frame's passes every frame): lets a caller keep graph set-up out of a
 * timed or latency-critical step call. */
int ausm_prepare_steps(ausm_ctx* ctx, int nsteps);
/* L(Q) of the current solution (boundary + fluxes + residual passes) */
int ausm_eval_residual(ausm_ctx* ctx, float* res);
/* per-cell wave speeds of the current solution (ausm.js:617-630) and their
 * max (cflReduction, ausm.js:535-582); ws may be NULL */
int ausm_wave_speeds(ausm_ctx* ctx, float* ws, float* max_ws);
/* surface force from row-1 pressure (graphs.js:16-101), summed in FP64:
 * out = (Fx, Fy, Cl, Cd); Cl and Cd are divided by q_inf * ref_length
 * (ref_length = 1 gives the reference's un-normalised values).  cp (M
 * values, optional) gets (p - p_inf)/q_inf per body segment. */
int ausm_forces(ausm_ctx* ctx, double ref_length, double* out, double* cp);
/* Derived display field of the current solution (the visualization pass,
 * src/shaders/ausm.js:586-710, up to its colour map): kind = the pass's
 * simDisplayMode, 0 schlieren exp(-2|grad rho|), 1 vorticity, 2 density,
 * 3 temperature p/rho, 4 pressure, 5 Mach (6: the same field, free-stream
 * value the inflow Mach instead of 0.5), 7 |v - V_inf| (the pass shows its
 * components), 8 entropy log(p/rho^gamma), 9 total pressure.  out: N x M
 * interior cells (row-major, i fastest); freestream_value (optional): the
 * value the pass's colour map is centred on. */
enum ausm_field {
  AUSM_FIELD_SCHLIEREN = 0, AUSM_FIELD_VORTICITY = 1, AUSM_FIELD_DENSITY = 2,
  AUSM_FIELD_TEMPERATURE = 3, AUSM_FIELD_PRESSURE = 4, AUSM_FIELD_MACH = 5,
  AUSM_FIELD_MACH_INFLOW = 6, AUSM_FIELD_VELOCITY_REL = 7, AUSM_FIELD_ENTROPY = 8,
  AUSM_FIELD_TOTAL_PRESSURE = 9
};
int ausm_derived_field(ausm_ctx* ctx, int kind, float* out, float* freestream_value);
/* FP64 L2 norm of the density residual of the current solution */
int ausm_residual_norm(ausm_ctx* ctx, double* l2);
/* FP64 state and residual, layouts as ausm_set_state / ausm_get_state /
 * ausm_eval_residual.  The state functions work in FP64 and FP32C precision
 * (FP32C: the value is split into / read back as hi + lo), the residual in
 * FP64 only. */
int ausm_set_state64(ausm_ctx* ctx, const double* q);
int ausm_get_state64(ausm_ctx* ctx, double* q);
int ausm_eval_residual64(ausm_ctx* ctx, double* res);

/* Self-check of the device arithmetic exact math relies on, no counterpart
 * in the reference.  The exact path divides with a refined hardware
 * reciprocal that must be the correctly rounded 1/d, and takes square roots
 * with the IEEE sequence stripped of its special-range handling.  Runs both
 * on `device` for every FP32 input of their ranges (divisors in
 * [2^-126, 2^126); square roots of +0 and of [2^-96, 2^128)) and returns the
 * number of inputs whose result differs from IEEE (0 on gfx950).  The
 * library keeps the count per device for the process: context and batch
 * creation and the first exact-math ausm_step_noalloc call on a device run
 * the check unless it has run, and all of them refuse a device that fails.
 * Callers that capture exact-math ausm_step_noalloc calls into a graph run
 * it once beforehand (the check cannot run under capture). */
int ausm_check_arithmetic(int device, long long* mismatches);

/* Timing hook, no counterpart in the reference (it times whole passes with
 * timestamp queries, main.js:597-602): launches each of the three stage
 * kernels `reps` times back to back on the context's stream between one
 * pair of HIP events and returns the mean duration per launch of each stage
 * in ms.  The solution is left untouched (stage outputs go to the scratch
 * buffers). */
int ausm_profile_stages(ausm_ctx* ctx, int reps, float* ms3);

/* ---- allocation-free step ------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc: nsteps TVD-RK3 steps of a caller-owned solution,
 * the same kernels and launch sequence as ausm_step on a context given the
 * same vertices, state and settings (bit-for-bit the same state), with no
 * context and no allocation.  vertices ((N+1) x M x 2) and state
 * ((N+3) x M x 4, updated in place, ghost rows rewritten) are DEVICE
 * pointers in the reference's layouts on `device`; stream is a hipStream_t
 * of that device (NULL: the null stream).  Every intermediate buffer comes
 * from one fixed device array of AUSM_NOALLOC_SCRATCH_MB MiB per GPU, held
 * until the process ends: grids needing more (about 1.2 million cells) are
 * AUSM_ERR_ARG.  All calls on one GPU share that array, so calls that
 * overlap on one GPU race.  Plain FP32 (the reference's arithmetic) only;
 * the first step takes its dt from the input state, later ones refresh it
 * every steps_per_dt_refresh steps.  info (optional) is read back with a
 * stream synchronisation; with info NULL the call only enqueues work,
 * except that the first exact-math call on a device runs the device's
 * arithmetic self-check (ausm_check_arithmetic) on `stream` and waits for
 * it, and refuses to (AUSM_ERR_ORDER) if the stream is capturing. */
#define AUSM_NOALLOC_SCRATCH_MB 64
typedef struct ausm_step_desc {
  int M, N;                   /* cells around / radial */
  double rho, p, vx, vy, gamma;  /* free stream, as ausm_set_freestream */
  int scheme, limiter;        /* as ausm_set_scheme */
  float K_p, K_u;
  float cfl, maxdt;           /* as ausm_set_cfl */
  int steps_per_dt_refresh;
  int math;                   /* AUSM_MATH_* */
  int dt_mode;                /* AUSM_DT_* */
  int rows_per_wave;          /* stage-kernel tile rows: 1, 2, 4, 8 (0 = 2) */
} ausm_step_desc;
int ausm_step_noalloc(int device, const ausm_step_desc* desc, const float* vertices, float* state,
                      int nsteps, void* stream, ausm_step_info* info);

/* ---- batched solves and continuous-batching service ---------------------
 * No counterpart in the reference (it runs one grid).  A batch context
 * holds many independent solves of different sizes resident in HBM and
 * advances all active ones together with one persistent kernel per RK
 * stage.  Descriptor fields mirror the single-instance setters above. */
typedef struct ausm_solve_desc {
  int M, N;              /* cells around / radial */
  const float* vertices; /* (N+1) x M x 2, FP32 (e.g. from ausm_grid_get_vertices) */
  double rho, p, vx, vy, gamma;
  int scheme, limiter;
  float K_p, K_u;
  float cfl, maxdt;
  int init_mode;         /* AUSM_INIT_* */
  int max_steps;         /* stop after this many steps */
  double res_drop;       /* converged when res_l2 <= res_drop * first res_l2 (0 = never) */
  double ref_length;     /* for Cl/Cd */
  int dt_mode;           /* AUSM_DT_GLOBAL or AUSM_DT_LOCAL (ausm_set_dt_mode) */
  int freeze_after;      /* van Albada limiters: freeze the limiter (ausm_freeze_limiter)
                            at the first convergence check at or after this many
                            steps; 0 = never */
} ausm_solve_desc;

typedef struct ausm_solve_status {
  int state;        /* 0 = queued, 1 = running, 2 = converged, 3 = step cap, 4 = failed */
  long long steps;  /* steps taken */
  double res_l2;    /* last density-residual norm */
  double res_l2_0;  /* first density-residual norm */
  double cl, cd;    /* forces at completion (or at the last poll) */
} ausm_solve_status;

typedef struct ausm_batch ausm_batch;

/* max_solves bounds the slots; hbm_budget_bytes (0 = free memory less a
 * reserve) bounds admission by device memory. */
int ausm_batch_create(int device, int max_solves, size_t hbm_budget_bytes, ausm_batch** out);
int ausm_batch_destroy(ausm_batch* b);
int ausm_batch_stream(ausm_batch* b, void** stream);
int ausm_batch_set_math(ausm_batch* b, int math);
/* AUSM_PREC_FP32 (default) or AUSM_PREC_FP32C (compensated state, see
 * ausm_set_precision) for every solve of the batch; only before the first
 * submit (or once the batch is empty) */
int ausm_batch_set_precision(ausm_batch* b, int precision);
/* fix the persistent schedule: rows per wavefront tile (1, 2, 4) and
 * workgroups per CU (1..16); 0 leaves a value to the per-mix autotuner */
int ausm_batch_set_schedule(ausm_batch* b, int rows_per_wave, int blocks_per_cu);
/* queue a solve; it is admitted when a slot and memory are free */
int ausm_batch_submit(ausm_batch* b, const ausm_solve_desc* desc, uint64_t* id);
/* advance every running solve by `nsteps` steps, rounded up to a multiple
 * of `check_every` (one persistent launch per RK stage per step); after every
 * `check_every` steps convergence is checked, finished solves retire and
 * queued ones are admitted.  *pending gets the number running or queued.
 * Step caps take effect at the first check at or after max_steps. */
int ausm_batch_advance(ausm_batch* b, int nsteps, int check_every, int* pending);
int ausm_batch_poll(ausm_batch* b, uint64_t id, ausm_solve_status* st);
/* copy a solve's state (finished or running) in the reference layout */
int ausm_batch_get_state(ausm_batch* b, uint64_t id, float* q);
int ausm_batch_release(ausm_batch* b, uint64_t id);

#ifdef __cplusplus
}
#endif
#endif /* AUSM_H */
