/*
 * Node-API binding of libausm (include/ausm.h) for Node.js hosts.
 *
 * Exposes the C ABI one-to-one to JavaScript: a context is an external
 * value; arrays cross as typed arrays (Float32Array / Float64Array) that are
 * copied by the library; errors become JS exceptions carrying
 * ausm_last_error().  stepAsync() runs the steps on the libuv thread pool
 * (napi_async_work) and resolves a Promise, so long step loops do not block
 * the event loop.
 */
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ausm.h"

#define NAPI_CALL(env, call)                                  \
  do {                                                        \
    if ((call) != napi_ok) {                                  \
      napi_throw_error((env), NULL, "N-API call failed: " #call); \
      return NULL;                                            \
    }                                                         \
  } while (0)

static napi_value throw_ausm(napi_env env, int rc) {
  char msg[600];
  snprintf(msg, sizeof msg, "ausm error %d: %s", rc, ausm_last_error());
  napi_throw_error(env, NULL, msg);
  return NULL;
}

#define AUSM_CALL(env, expr)               \
  do {                                     \
    int _rc = (expr);                      \
    if (_rc != AUSM_OK) return throw_ausm((env), _rc); \
  } while (0)

/* a context with the grid size it was created for: typed arrays crossing
 * the boundary are checked against it (the library reads and writes
 * arrays of exactly these sizes) */
typedef struct {
  ausm_ctx* c;
  int M, N;
} ctx_box;

static void ctx_finalize(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  ctx_box* b = (ctx_box*)data;
  if (b) {
    ausm_ctx_destroy(b->c);
    free(b);
  }
}

/* ---- argument helpers ---------------------------------------------------- */
static int get_args(napi_env env, napi_callback_info info, size_t n, napi_value* argv) {
  size_t argc = n;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return 0;
  if (argc < n) {
    napi_throw_type_error(env, NULL, "too few arguments");
    return 0;
  }
  return 1;
}

static double num(napi_env env, napi_value v) {
  double d = 0.0;
  napi_get_value_double(env, v, &d);
  return d;
}

static ctx_box* box_of(napi_env env, napi_value v) {
  void* p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, NULL, "expected an ausm context");
    return NULL;
  }
  return (ctx_box*)p;
}

static ausm_ctx* ctx_of(napi_env env, napi_value v) {
  ctx_box* b = box_of(env, v);
  return b ? b->c : NULL;
}

/* RangeError unless a typed array holds exactly `want` elements */
static int check_len(napi_env env, size_t got, size_t want, const char* what) {
  if (got == want) return 1;
  char msg[200];
  snprintf(msg, sizeof msg, "%s: expected %zu elements, got %zu", what, want, got);
  napi_throw_range_error(env, NULL, msg);
  return 0;
}

/* RangeError unless the (M, N) a caller passes are the context's own */
static int check_dims(napi_env env, const ctx_box* b, napi_value m, napi_value n) {
  const int M = (int)num(env, m), N = n ? (int)num(env, n) : b->N;
  if (M == b->M && N == b->N) return 1;
  char msg[200];
  snprintf(msg, sizeof msg, "grid size %d x %d does not match the context's %d x %d", M, N, b->M,
           b->N);
  napi_throw_range_error(env, NULL, msg);
  return 0;
}

/* pointer and element count of a typed array of the given type */
static void* typed(napi_env env, napi_value v, napi_typedarray_type want, size_t* n) {
  napi_typedarray_type t;
  size_t len = 0;
  void* data = NULL;
  napi_value ab;
  size_t off = 0;
  if (napi_get_typedarray_info(env, v, &t, &len, &data, &ab, &off) != napi_ok || t != want) {
    napi_throw_type_error(env, NULL,
                          want == napi_float32_array ? "expected a Float32Array" : "expected a Float64Array");
    return NULL;
  }
  *n = len;
  return data;
}

static napi_value new_typed(napi_env env, napi_typedarray_type t, size_t n, void** data) {
  napi_value ab, arr;
  const size_t esz = t == napi_float64_array ? 8 : 4;
  if (napi_create_arraybuffer(env, n * esz, data, &ab) != napi_ok) return NULL;
  if (napi_create_typedarray(env, t, n, ab, 0, &arr) != napi_ok) return NULL;
  return arr;
}

static void set_num(napi_env env, napi_value obj, const char* k, double v) {
  napi_value x;
  napi_create_double(env, v, &x);
  napi_set_named_property(env, obj, k, x);
}

/* ---- context --------------------------------------------------------------- */
static napi_value js_create(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return NULL;
  ausm_ctx* c = NULL;
  const int M = (int)num(env, a[0]), N = (int)num(env, a[1]);
  AUSM_CALL(env, ausm_ctx_create((int)num(env, a[2]), M, N, &c));
  ctx_box* b = (ctx_box*)malloc(sizeof(ctx_box));
  if (!b) {
    ausm_ctx_destroy(c);
    return throw_ausm(env, AUSM_ERR_NOMEM);
  }
  b->c = c;
  b->M = M;
  b->N = N;
  napi_value ext;
  if (napi_create_external(env, b, ctx_finalize, NULL, &ext) != napi_ok) {
    ctx_finalize(env, b, NULL);
    napi_throw_error(env, NULL, "N-API call failed: napi_create_external");
    return NULL;
  }
  return ext;
}

/* ---- geometry ---------------------------------------------------------------- */
static napi_value js_naca4(napi_env env, napi_callback_info info) {
  napi_value a[8];
  if (!get_args(env, info, 8, a)) return NULL;
  const int M = (int)num(env, a[7]);
  if (M < 4) return throw_ausm(env, AUSM_ERR_ARG);
  void* d;
  napi_value out = new_typed(env, napi_float64_array, 2 * (size_t)M, &d);
  if (!out) return NULL;
  AUSM_CALL(env, ausm_boundary_naca4(num(env, a[0]), num(env, a[1]), num(env, a[2]), num(env, a[3]),
                                     num(env, a[4]), num(env, a[5]), (int)num(env, a[6]), M, (double*)d));
  return out;
}

static napi_value js_circle(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return NULL;
  const int M = (int)num(env, a[1]);
  if (M < 4) return throw_ausm(env, AUSM_ERR_ARG);
  void* d;
  napi_value out = new_typed(env, napi_float64_array, 2 * (size_t)M, &d);
  if (!out) return NULL;
  AUSM_CALL(env, ausm_boundary_circle(num(env, a[0]), M, (double*)d));
  return out;
}

static napi_value js_rect(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return NULL;
  const int M = (int)num(env, a[1]);
  if (M < 4) return throw_ausm(env, AUSM_ERR_ARG);
  void* d;
  napi_value out = new_typed(env, napi_float64_array, 2 * (size_t)M, &d);
  if (!out) return NULL;
  AUSM_CALL(env, ausm_boundary_rect(num(env, a[0]), M, (double*)d));
  return out;
}

static napi_value js_cst(napi_env env, napi_callback_info info) {
  napi_value a[8];
  if (!get_args(env, info, 8, a)) return NULL;
  size_t nu = 0, nl = 0;
  const double* au = (const double*)typed(env, a[0], napi_float64_array, &nu);
  if (!au) return NULL;
  const double* al = (const double*)typed(env, a[1], napi_float64_array, &nl);
  if (!al) return NULL;
  if (nu != nl) return throw_ausm(env, AUSM_ERR_ARG);
  const int M = (int)num(env, a[7]);
  if (M < 4) return throw_ausm(env, AUSM_ERR_ARG);
  void* d;
  napi_value out = new_typed(env, napi_float64_array, 2 * (size_t)M, &d);
  if (!out) return NULL;
  AUSM_CALL(env, ausm_boundary_cst(au, al, (int)nu, num(env, a[2]), num(env, a[3]), num(env, a[4]),
                                   num(env, a[5]), (int)num(env, a[6]), M, (double*)d));
  return out;
}

static napi_value js_bezier(napi_env env, napi_callback_info info) {
  napi_value a[7];
  if (!get_args(env, info, 7, a)) return NULL;
  size_t nu = 0, nl = 0;
  const double* cu = (const double*)typed(env, a[0], napi_float64_array, &nu);
  if (!cu) return NULL;
  const double* cl = (const double*)typed(env, a[1], napi_float64_array, &nl);
  if (!cl) return NULL;
  if (nu % 2 || nl % 2) return throw_ausm(env, AUSM_ERR_ARG);
  const int M = (int)num(env, a[6]);
  if (M < 4) return throw_ausm(env, AUSM_ERR_ARG);
  void* d;
  napi_value out = new_typed(env, napi_float64_array, 2 * (size_t)M, &d);
  if (!out) return NULL;
  AUSM_CALL(env, ausm_boundary_bezier(cu, (int)(nu / 2), cl, (int)(nl / 2), num(env, a[2]),
                                      num(env, a[3]), num(env, a[4]), (int)num(env, a[5]), M,
                                      (double*)d));
  return out;
}

static napi_value js_sears_haack(napi_env env, napi_callback_info info) {
  napi_value a[4];
  if (!get_args(env, info, 4, a)) return NULL;
  const int M = (int)num(env, a[3]);
  if (M < 4) return throw_ausm(env, AUSM_ERR_ARG);
  void* d;
  napi_value out = new_typed(env, napi_float64_array, 2 * (size_t)M, &d);
  if (!out) return NULL;
  AUSM_CALL(env, ausm_boundary_sears_haack(num(env, a[0]), num(env, a[1]), num(env, a[2]), M, (double*)d));
  return out;
}

static napi_value js_haack(napi_env env, napi_callback_info info) {
  napi_value a[8];
  if (!get_args(env, info, 8, a)) return NULL;
  const int M = (int)num(env, a[7]);
  if (M < 4) return throw_ausm(env, AUSM_ERR_ARG);
  void* d;
  napi_value out = new_typed(env, napi_float64_array, 2 * (size_t)M, &d);
  if (!out) return NULL;
  AUSM_CALL(env, ausm_boundary_haack_projectile(num(env, a[0]), num(env, a[1]), num(env, a[2]),
                                                num(env, a[3]), num(env, a[4]), num(env, a[5]),
                                                num(env, a[6]), M, (double*)d));
  return out;
}

static napi_value js_polygon(napi_env env, napi_callback_info info) {
  napi_value a[6];
  if (!get_args(env, info, 6, a)) return NULL;
  const int M = (int)num(env, a[5]);
  if (M < 4) return throw_ausm(env, AUSM_ERR_ARG);
  void* d;
  napi_value out = new_typed(env, napi_float64_array, 2 * (size_t)M, &d);
  if (!out) return NULL;
  AUSM_CALL(env, ausm_boundary_polygon(num(env, a[0]), (int)num(env, a[1]), num(env, a[2]),
                                       num(env, a[3]), num(env, a[4]), M, (double*)d));
  return out;
}

/* boundaryDat(text, size, xOffset, M) -> Float64Array (2M) */
static napi_value js_dat(napi_env env, napi_callback_info info) {
  napi_value a[4];
  if (!get_args(env, info, 4, a)) return NULL;
  size_t len = 0;
  NAPI_CALL(env, napi_get_value_string_utf8(env, a[0], NULL, 0, &len));
  char* text = (char*)malloc(len + 1);
  if (!text) return throw_ausm(env, AUSM_ERR_NOMEM);
  if (napi_get_value_string_utf8(env, a[0], text, len + 1, &len) != napi_ok) {
    free(text);
    napi_throw_error(env, NULL, "expected a string");
    return NULL;
  }
  const int M = (int)num(env, a[3]);
  if (M < 4) {
    free(text);
    return throw_ausm(env, AUSM_ERR_ARG);
  }
  void* d;
  napi_value out = new_typed(env, napi_float64_array, 2 * (size_t)M, &d);
  if (!out) {
    free(text);
    return NULL;
  }
  const int rc = ausm_boundary_dat(text, num(env, a[1]), num(env, a[2]), M, (double*)d, NULL);
  free(text);
  if (rc != AUSM_OK) return throw_ausm(env, rc);
  return out;
}

/* ---- grid -------------------------------------------------------------------- */
static napi_value js_grid_from_boundary(napi_env env, napi_callback_info info) {
  napi_value a[6];
  if (!get_args(env, info, 6, a)) return NULL;
  ctx_box* b = box_of(env, a[0]);
  if (!b) return NULL;
  size_t nb, no;
  double* body = (double*)typed(env, a[1], napi_float64_array, &nb);
  if (!body || !check_len(env, nb, 2 * (size_t)b->M, "body points")) return NULL;
  double* outer = (double*)typed(env, a[2], napi_float64_array, &no);
  if (!outer || !check_len(env, no, 2 * (size_t)b->M, "far-field points")) return NULL;
  ausm_ctx* c = b->c;
  AUSM_CALL(env, ausm_grid_from_boundary(c, body, outer, (int)num(env, a[3]), (int)num(env, a[4]),
                                         num(env, a[5])));
  return NULL;
}

static napi_value js_set_vertices(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return NULL;
  ctx_box* b = box_of(env, a[0]);
  if (!b) return NULL;
  size_t n;
  float* v = (float*)typed(env, a[1], napi_float32_array, &n);
  if (!v || !check_len(env, n, 2 * (size_t)b->M * ((size_t)b->N + 1), "vertices")) return NULL;
  AUSM_CALL(env, ausm_grid_set_vertices(b->c, v));
  return NULL;
}

static napi_value js_get_vertices(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return NULL;
  ctx_box* b = box_of(env, a[0]);
  if (!b || !check_dims(env, b, a[1], a[2])) return NULL;
  const size_t n = (size_t)b->M * ((size_t)b->N + 1) * 2;
  void* d;
  napi_value out = new_typed(env, napi_float32_array, n, &d);
  if (!out) return NULL;
  AUSM_CALL(env, ausm_grid_get_vertices(b->c, (float*)d));
  return out;
}

/* ---- parameters ------------------------------------------------------------------ */
static napi_value js_set_freestream(napi_env env, napi_callback_info info) {
  napi_value a[6];
  if (!get_args(env, info, 6, a)) return NULL;
  ausm_ctx* c = ctx_of(env, a[0]);
  if (!c) return NULL;
  AUSM_CALL(env, ausm_set_freestream(c, num(env, a[1]), num(env, a[2]), num(env, a[3]), num(env, a[4]),
                                     num(env, a[5])));
  return NULL;
}

static napi_value js_set_scheme(napi_env env, napi_callback_info info) {
  napi_value a[5];
  if (!get_args(env, info, 5, a)) return NULL;
  ausm_ctx* c = ctx_of(env, a[0]);
  if (!c) return NULL;
  AUSM_CALL(env, ausm_set_scheme(c, (int)num(env, a[1]), (int)num(env, a[2]), (float)num(env, a[3]),
                                 (float)num(env, a[4])));
  return NULL;
}

static napi_value js_set_cfl(napi_env env, napi_callback_info info) {
  napi_value a[4];
  if (!get_args(env, info, 4, a)) return NULL;
  ausm_ctx* c = ctx_of(env, a[0]);
  if (!c) return NULL;
  AUSM_CALL(env, ausm_set_cfl(c, (float)num(env, a[1]), (float)num(env, a[2]), (int)num(env, a[3])));
  return NULL;
}

static napi_value js_set_math(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return NULL;
  ausm_ctx* c = ctx_of(env, a[0]);
  if (!c) return NULL;
  AUSM_CALL(env, ausm_set_math(c, (int)num(env, a[1])));
  return NULL;
}

/* ---- state --------------------------------------------------------------------- */
static napi_value js_init_state(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return NULL;
  ausm_ctx* c = ctx_of(env, a[0]);
  if (!c) return NULL;
  AUSM_CALL(env, ausm_init_state(c, (int)num(env, a[1])));
  return NULL;
}

static napi_value js_set_state(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return NULL;
  ctx_box* b = box_of(env, a[0]);
  if (!b) return NULL;
  size_t n;
  float* q = (float*)typed(env, a[1], napi_float32_array, &n);
  if (!q || !check_len(env, n, 4 * (size_t)b->M * ((size_t)b->N + 3), "state")) return NULL;
  AUSM_CALL(env, ausm_set_state(b->c, q));
  return NULL;
}

static napi_value js_get_state(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return NULL;
  ctx_box* b = box_of(env, a[0]);
  if (!b || !check_dims(env, b, a[1], a[2])) return NULL;
  const size_t n = (size_t)b->M * ((size_t)b->N + 3) * 4;
  void* d;
  napi_value out = new_typed(env, napi_float32_array, n, &d);
  if (!out) return NULL;
  AUSM_CALL(env, ausm_get_state(b->c, (float*)d));
  return out;
}

static napi_value info_object(napi_env env, const ausm_step_info* si) {
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "dt", si->dt);
  set_num(env, o, "maxWaveSpeed", si->max_wave_speed);
  set_num(env, o, "resL2", si->res_l2);
  set_num(env, o, "nonfinite", si->nonfinite);
  set_num(env, o, "stepsDone", (double)si->steps_done);
  return o;
}

static napi_value js_step(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return NULL;
  ausm_ctx* c = ctx_of(env, a[0]);
  if (!c) return NULL;
  ausm_step_info si;
  memset(&si, 0, sizeof si);
  AUSM_CALL(env, ausm_step(c, (int)num(env, a[1]), &si));
  return info_object(env, &si);
}

/* asynchronous steps on the libuv pool */
typedef struct {
  ausm_ctx* ctx;
  int n;
  int rc;
  char err[512];
  ausm_step_info si;
  napi_deferred deferred;
  napi_async_work work;
} step_job;

static void step_execute(napi_env env, void* data) {
  (void)env;
  step_job* j = (step_job*)data;
  memset(&j->si, 0, sizeof j->si);
  j->rc = ausm_step(j->ctx, j->n, &j->si);
  if (j->rc != AUSM_OK) snprintf(j->err, sizeof j->err, "ausm error %d: %s", j->rc, ausm_last_error());
}

static void step_complete(napi_env env, napi_status status, void* data) {
  step_job* j = (step_job*)data;
  if (status == napi_ok && j->rc == AUSM_OK) {
    napi_resolve_deferred(env, j->deferred, info_object(env, &j->si));
  } else {
    napi_value msg, err;
    napi_create_string_utf8(env, j->rc != AUSM_OK ? j->err : "step cancelled", NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, NULL, msg, &err);
    napi_reject_deferred(env, j->deferred, err);
  }
  napi_delete_async_work(env, j->work);
  free(j);
}

static napi_value js_step_async(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return NULL;
  ausm_ctx* c = ctx_of(env, a[0]);
  if (!c) return NULL;
  step_job* j = (step_job*)calloc(1, sizeof(step_job));
  j->ctx = c;
  j->n = (int)num(env, a[1]);
  napi_value promise, name;
  NAPI_CALL(env, napi_create_promise(env, &j->deferred, &promise));
  NAPI_CALL(env, napi_create_string_utf8(env, "ausm_step", NAPI_AUTO_LENGTH, &name));
  NAPI_CALL(env, napi_create_async_work(env, NULL, name, step_execute, step_complete, j, &j->work));
  NAPI_CALL(env, napi_queue_async_work(env, j->work));
  return promise;
}

static napi_value js_forces(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return NULL;
  ctx_box* b = box_of(env, a[0]);
  if (!b || !check_dims(env, b, a[2], NULL)) return NULL;
  double out[4];
  void* cp;
  napi_value cpa = new_typed(env, napi_float64_array, (size_t)b->M, &cp);
  if (!cpa) return NULL;
  AUSM_CALL(env, ausm_forces(b->c, num(env, a[1]), out, (double*)cp));
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "fx", out[0]);
  set_num(env, o, "fy", out[1]);
  set_num(env, o, "cl", out[2]);
  set_num(env, o, "cd", out[3]);
  napi_set_named_property(env, o, "cp", cpa);
  return o;
}

static napi_value js_residual(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return NULL;
  ctx_box* b = box_of(env, a[0]);
  if (!b || !check_dims(env, b, a[1], a[2])) return NULL;
  const size_t n = (size_t)b->M * (size_t)b->N * 4;
  void* d;
  napi_value out = new_typed(env, napi_float32_array, n, &d);
  if (!out) return NULL;
  AUSM_CALL(env, ausm_eval_residual(b->c, (float*)d));
  return out;
}

static napi_value js_set_precision(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return NULL;
  ausm_ctx* c = ctx_of(env, a[0]);
  if (!c) return NULL;
  AUSM_CALL(env, ausm_set_precision(c, (int)num(env, a[1])));
  return NULL;
}

static napi_value js_set_dt_mode(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return NULL;
  ausm_ctx* c = ctx_of(env, a[0]);
  if (!c) return NULL;
  AUSM_CALL(env, ausm_set_dt_mode(c, (int)num(env, a[1])));
  return NULL;
}

/* derivedField(ctx, kind, M, N) -> {field: Float32Array (N*M), freestream} */
static napi_value js_derived_field(napi_env env, napi_callback_info info) {
  napi_value a[4];
  if (!get_args(env, info, 4, a)) return NULL;
  ctx_box* b = box_of(env, a[0]);
  if (!b || !check_dims(env, b, a[2], a[3])) return NULL;
  const size_t n = (size_t)b->M * (size_t)b->N;
  void* d;
  napi_value arr = new_typed(env, napi_float32_array, n, &d);
  if (!arr) return NULL;
  float fs = 0.0f;
  AUSM_CALL(env, ausm_derived_field(b->c, (int)num(env, a[1]), (float*)d, &fs));
  napi_value obj;
  NAPI_CALL(env, napi_create_object(env, &obj));
  napi_set_named_property(env, obj, "field", arr);
  set_num(env, obj, "freestream", fs);
  return obj;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
  (void)env;
  (void)info;
  /* contexts are released by the external's finalizer */
  return NULL;
}

static napi_value js_version(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value v;
  napi_create_int32(env, ausm_version(), &v);
  return v;
}

static napi_value init(napi_env env, napi_value exports) {
  const napi_property_descriptor props[] = {
      {"version", NULL, js_version, NULL, NULL, NULL, napi_default, NULL},
      {"createContext", NULL, js_create, NULL, NULL, NULL, napi_default, NULL},
      {"destroyContext", NULL, js_destroy, NULL, NULL, NULL, napi_default, NULL},
      {"boundaryNaca4", NULL, js_naca4, NULL, NULL, NULL, napi_default, NULL},
      {"boundaryCircle", NULL, js_circle, NULL, NULL, NULL, napi_default, NULL},
      {"boundaryRect", NULL, js_rect, NULL, NULL, NULL, napi_default, NULL},
      {"boundaryCst", NULL, js_cst, NULL, NULL, NULL, napi_default, NULL},
      {"boundaryBezier", NULL, js_bezier, NULL, NULL, NULL, napi_default, NULL},
      {"boundarySearsHaack", NULL, js_sears_haack, NULL, NULL, NULL, napi_default, NULL},
      {"boundaryHaackProjectile", NULL, js_haack, NULL, NULL, NULL, napi_default, NULL},
      {"boundaryPolygon", NULL, js_polygon, NULL, NULL, NULL, napi_default, NULL},
      {"boundaryDat", NULL, js_dat, NULL, NULL, NULL, napi_default, NULL},
      {"setPrecision", NULL, js_set_precision, NULL, NULL, NULL, napi_default, NULL},
      {"setDtMode", NULL, js_set_dt_mode, NULL, NULL, NULL, napi_default, NULL},
      {"derivedField", NULL, js_derived_field, NULL, NULL, NULL, napi_default, NULL},
      {"gridFromBoundary", NULL, js_grid_from_boundary, NULL, NULL, NULL, napi_default, NULL},
      {"setVertices", NULL, js_set_vertices, NULL, NULL, NULL, napi_default, NULL},
      {"getVertices", NULL, js_get_vertices, NULL, NULL, NULL, napi_default, NULL},
      {"setFreestream", NULL, js_set_freestream, NULL, NULL, NULL, napi_default, NULL},
      {"setScheme", NULL, js_set_scheme, NULL, NULL, NULL, napi_default, NULL},
      {"setCfl", NULL, js_set_cfl, NULL, NULL, NULL, napi_default, NULL},
      {"setMath", NULL, js_set_math, NULL, NULL, NULL, napi_default, NULL},
      {"initState", NULL, js_init_state, NULL, NULL, NULL, napi_default, NULL},
      {"setState", NULL, js_set_state, NULL, NULL, NULL, napi_default, NULL},
      {"getState", NULL, js_get_state, NULL, NULL, NULL, napi_default, NULL},
      {"step", NULL, js_step, NULL, NULL, NULL, napi_default, NULL},
      {"stepAsync", NULL, js_step_async, NULL, NULL, NULL, napi_default, NULL},
      {"forces", NULL, js_forces, NULL, NULL, NULL, napi_default, NULL},
      {"residual", NULL, js_residual, NULL, NULL, NULL, napi_default, NULL},
  };
  napi_define_properties(env, exports, sizeof props / sizeof props[0], props);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
