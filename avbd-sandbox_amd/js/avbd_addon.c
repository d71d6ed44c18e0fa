/*
 * avbd_addon.c — Node-API binding of include/avbd.h for the JavaScript host.
 *
 * Thin by design: each export calls one C-ABI function, arrays cross as typed
 * arrays (Float64Array / Int32Array / Uint32Array) and are copied by the
 * library. sync() runs on the libuv thread pool and returns a Promise, like
 * GpuBackend.idle() (backend.js:1804-1818). Errors become thrown Errors
 * carrying avbd_last_error().
 */
#include <math.h>
#include <stdio.h>
#include <node_api.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/avbd.h"

#define CHECK(env, call)                                       \
  do {                                                         \
    if ((call) != napi_ok) {                                   \
      napi_throw_error((env), NULL, "N-API call failed: " #call); \
      return NULL;                                             \
    }                                                          \
  } while (0)

/* Throws an Error whose .code is the AVBD_E_* value as a string ("-3"). */
static napi_value throw_avbd(napi_env env, avbd_ctx *ctx, int rc) {
  char msg[512], code[16];
  snprintf(msg, sizeof msg, "avbd error %d: %s", rc, ctx ? avbd_last_error(ctx) : "no context");
  snprintf(code, sizeof code, "%d", rc);
  napi_throw_error(env, code, msg);
  return NULL;
}

/* The context external: the library context and the number of sync() jobs
 * in flight on the thread pool. While one runs, every other call on the
 * context throws (the library's calls are not re-entrant, and a pack would
 * destroy the stream the job is waiting on). */
typedef struct {
  avbd_ctx *ctx;
  int syncs;
} Handle;

static Handle *get_handle(napi_env env, napi_value v) {
  void *p = NULL;
  if (!v || napi_get_value_external(env, v, &p) != napi_ok) return NULL;
  return (Handle *)p;
}

/* The context of argument v for a call that is not sync(); NULL after
 * throwing when v is not a context or a sync is in flight. */
static avbd_ctx *get_ctx(napi_env env, napi_value v) {
  Handle *h = get_handle(env, v);
  if (!h || !h->ctx) {
    napi_throw_type_error(env, NULL, "the first argument must be a context from create()");
    return NULL;
  }
  if (h->syncs > 0) {
    napi_throw_error(env, NULL, "a sync() of this context is in flight: await it before any other call");
    return NULL;
  }
  return h->ctx;
}

static const char *type_name(napi_typedarray_type t) {
  switch (t) {
    case napi_int32_array: return "Int32Array";
    case napi_uint32_array: return "Uint32Array";
    case napi_float32_array: return "Float32Array";
    case napi_float64_array: return "Float64Array";
    default: return "another typed array";
  }
}

/* Argument v as a typed array of element type `type` holding exactly `want`
 * elements (want < 0: any count, returned in *count). An optional argument may
 * be absent, null or undefined (*data = NULL). Returns 1, or 0 after throwing
 * a TypeError (not a typed array of that type) or a RangeError (wrong length),
 * so that no array reaches the library with fewer elements than it reads or
 * writes. */
static int arr_arg(napi_env env, napi_value v, napi_typedarray_type type, const char *name, long long want,
                   int optional, void **data, size_t *count) {
  char msg[160];
  *data = NULL;
  if (count) *count = 0;
  napi_valuetype vt = napi_undefined;
  if (v && napi_typeof(env, v, &vt) != napi_ok) vt = napi_undefined;
  if (!v || vt == napi_undefined || vt == napi_null) {
    if (optional) return 1;
    snprintf(msg, sizeof msg, "%s is required (a %s)", name, type_name(type));
    napi_throw_type_error(env, NULL, msg);
    return 0;
  }
  bool is = false;
  napi_typedarray_type t;
  size_t n = 0, off = 0;
  void *p = NULL;
  napi_value ab;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is ||
      napi_get_typedarray_info(env, v, &t, &n, &p, &ab, &off) != napi_ok || t != type) {
    snprintf(msg, sizeof msg, "%s must be a %s", name, type_name(type));
    napi_throw_type_error(env, NULL, msg);
    return 0;
  }
  if (want >= 0 && (long long)n != want) {
    snprintf(msg, sizeof msg, "%s holds %zu elements, %lld expected", name, n, want);
    napi_throw_range_error(env, NULL, msg);
    return 0;
  }
  *data = p;
  if (count) *count = n;
  return 1;
}

static double get_num(napi_env env, napi_value obj, const char *key, double dflt) {
  napi_value v;
  bool has = false;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return dflt;
  if (napi_get_named_property(env, obj, key, &v) != napi_ok) return dflt;
  double d;
  if (napi_get_value_double(env, v, &d) != napi_ok) return dflt;
  return d;
}

static napi_value get_prop(napi_env env, napi_value obj, const char *key) {
  napi_value v = NULL;
  bool has = false;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return NULL;
  napi_get_named_property(env, obj, key, &v);
  napi_valuetype t;
  napi_typeof(env, v, &t);
  return (t == napi_undefined || t == napi_null) ? NULL : v;
}

static void finalize_ctx(napi_env env, void *data, void *hint) {
  (void)env;
  (void)hint;
  Handle *h = (Handle *)data;
  if (h) {
    if (h->ctx) avbd_destroy(h->ctx);
    free(h);
  }
}

/* create({device, precision, team, manifoldCap, pairCap, sceneSolve, graphs}) -> external (GpuBackend.init, backend.js:337-469) */
static napi_value js_create(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_config cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.precision = AVBD_PREC_MIXED;
  if (argc > 0) {
    cfg.device = (int32_t)get_num(env, argv[0], "device", 0);
    cfg.precision = (int32_t)get_num(env, argv[0], "precision", AVBD_PREC_MIXED);
    cfg.team = (int32_t)get_num(env, argv[0], "team", 0);
    cfg.manifold_cap = (int32_t)get_num(env, argv[0], "manifoldCap", 0);
    cfg.pair_cap = (int32_t)get_num(env, argv[0], "pairCap", 0);
    cfg.scene_solve = (int32_t)get_num(env, argv[0], "sceneSolve", 0);
    cfg.graphs = (int32_t)get_num(env, argv[0], "graphs", 0);
  }
  avbd_ctx *ctx = NULL;
  int rc = avbd_create(&ctx, &cfg);
  if (rc) return throw_avbd(env, NULL, rc);
  Handle *h = (Handle *)calloc(1, sizeof(Handle));
  if (!h) {
    avbd_destroy(ctx);
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  h->ctx = ctx;
  napi_value ext;
  if (napi_create_external(env, h, finalize_ctx, NULL, &ext) != napi_ok) {
    avbd_destroy(ctx);
    free(h);
    napi_throw_error(env, NULL, "N-API call failed: napi_create_external");
    return NULL;
  }
  return ext;
}

/* setParams(ctx, {dt, gravity, iterations, ...}) — Solver.defaultParams names (solver.js:92-236) */
static napi_value js_set_params(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2] = {NULL, NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  napi_valuetype t = napi_undefined;
  if (argc < 2 || napi_typeof(env, argv[1], &t) != napi_ok || t != napi_object) {
    napi_throw_type_error(env, NULL, "params must be an object");
    return NULL;
  }
  avbd_params p;
  avbd_default_params(&p);
  napi_value o = argv[1];
  p.dt = get_num(env, o, "dt", p.dt);
  p.gravity = get_num(env, o, "gravity", p.gravity);
  p.beta_lin = get_num(env, o, "betaLin", p.beta_lin);
  p.beta_ang = get_num(env, o, "betaAng", p.beta_ang);
  p.alpha = get_num(env, o, "alpha", p.alpha);
  p.gamma = get_num(env, o, "gamma", p.gamma);
  p.penalty_min = get_num(env, o, "penaltyMin", p.penalty_min);
  p.penalty_max = get_num(env, o, "penaltyMax", p.penalty_max);
  p.collision_margin = get_num(env, o, "collisionMargin", p.collision_margin);
  p.stick_threshold = get_num(env, o, "stickThreshold", p.stick_threshold);
  p.iterations = (int32_t)get_num(env, o, "iterations", p.iterations);
  p.post_stabilize = (int32_t)get_num(env, o, "postStabilize", p.post_stabilize);
  p.rotated_inertia = (int32_t)get_num(env, o, "rotatedInertia", p.rotated_inertia);
  p.cached_contact_jacobians = (int32_t)get_num(env, o, "cachedContactJacobians", p.cached_contact_jacobians);
  p.color_order = (int32_t)get_num(env, o, "colorOrder", p.color_order);
  /* the GPU contact law (solver.js:142-162) */
  p.contact_model = (int32_t)get_num(env, o, "contactModel", p.contact_model);
  p.contact_offset = get_num(env, o, "contactOffset", p.contact_offset);
  p.rest_offset = get_num(env, o, "restOffset", p.rest_offset);
  p.restitution_threshold = get_num(env, o, "restitutionThreshold", p.restitution_threshold);
  p.contact_persistence = get_num(env, o, "contactPersistence", p.contact_persistence);
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int rc = avbd_set_params(ctx, &p);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* pack(ctx, arenas[, oldIndex]) with typed arrays named as avbd_scene's
 * fields (GpuBackend.pack, backend.js:662-1301). The body count is the length
 * of `shape`; every other body array must hold exactly that many records, and
 * every joint array as many as `jointA`. */
static napi_value js_pack(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3] = {NULL, NULL, NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  napi_valuetype t = napi_undefined;
  if (argc < 2 || napi_typeof(env, argv[1], &t) != napi_ok || t != napi_object) {
    napi_throw_type_error(env, NULL, "the scene must be an object of typed arrays");
    return NULL;
  }
  napi_value o = argv[1];
  avbd_scene s;
  memset(&s, 0, sizeof s);
  size_t n = 0, nj = 0;
  void *d;
  const napi_typedarray_type I32 = napi_int32_array, F64 = napi_float64_array;
#define ARR(field, key, type, want, opt)                                            \
  if (!arr_arg(env, get_prop(env, o, key), type, key, (want), (opt), &d, NULL)) \
    return NULL;                                                                    \
  s.field = d;
  if (!arr_arg(env, get_prop(env, o, "shape"), I32, "shape", -1, 0, &d, &n)) return NULL;
  s.shape = d;
  const long long N = (long long)n;
  ARR(size, "size", F64, 3 * N, 0);
  ARR(density, "density", F64, N, 0);
  ARR(friction, "friction", F64, N, 0);
  ARR(pos, "pos", F64, 3 * N, 0);
  ARR(quat, "quat", F64, 4 * N, 1);
  ARR(vel, "vel", F64, 3 * N, 1);
  ARR(angvel, "angvel", F64, 3 * N, 1);
  ARR(scene_of, "sceneOf", I32, N, 1);
  /* the GPU contact law's material (Rigid.setMaterial, rigid.js:144-214) */
  ARR(static_friction, "staticFriction", F64, N, 1);
  ARR(dynamic_friction, "dynamicFriction", F64, N, 1);
  ARR(restitution, "restitution", F64, N, 1);
  ARR(contact_offset, "contactOffset", F64, N, 1);
  ARR(rest_offset, "restOffset", F64, N, 1);
  if (!arr_arg(env, get_prop(env, o, "jointA"), I32, "jointA", -1, 1, &d, &nj)) return NULL;
  s.joint_a = d;
  if (s.joint_a) {
    const long long NJ = (long long)nj;
    ARR(joint_b, "jointB", I32, NJ, 0);
    ARR(joint_ra, "jointRA", F64, 3 * NJ, 0);
    ARR(joint_rb, "jointRB", F64, 3 * NJ, 0);
    ARR(joint_stiff_lin, "jointStiffLin", F64, NJ, 0);
    ARR(joint_stiff_ang, "jointStiffAng", F64, NJ, 0);
    ARR(joint_fracture, "jointFracture", F64, NJ, 0);
    ARR(joint_kind, "jointKind", I32, NJ, 1);
  }
#undef ARR
  s.num_bodies = (int32_t)n;
  s.num_joints = s.joint_a ? (int32_t)nj : 0;
  /* pack(ctx, scene, oldIndex Int32Array): avbd_repack, keeping state */
  if (!arr_arg(env, argc > 2 ? argv[2] : NULL, I32, "oldIndex", N, 1, &d, NULL)) return NULL;
  const int32_t *old_index = (const int32_t *)d;
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int rc = old_index ? avbd_repack(ctx, &s, old_index) : avbd_pack(ctx, &s);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* step(ctx, n) (GpuBackend.step, backend.js:1393-1620) */
static napi_value js_step(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2] = {NULL, NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  int32_t n = 1;
  if (argc > 1) napi_get_value_int32(env, argv[1], &n);
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int rc = avbd_step(ctx, n);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* sync(ctx) -> Promise (GpuBackend.idle, backend.js:1804-1818). The job holds
 * a reference on the context external, so that it outlives a host that drops
 * the context meanwhile, and counts itself in the handle, so that every other
 * call on the context throws until it has completed. */
typedef struct {
  Handle *h;
  napi_ref ref;
  int rc;
  char err[448];
  napi_deferred deferred;
  napi_async_work work;
} SyncJob;

static void sync_execute(napi_env env, void *data) {
  (void)env;
  SyncJob *j = (SyncJob *)data;
  j->rc = avbd_sync(j->h->ctx);
  if (j->rc) snprintf(j->err, sizeof j->err, "%s", avbd_last_error(j->h->ctx));
}

static void sync_complete(napi_env env, napi_status status, void *data) {
  SyncJob *j = (SyncJob *)data;
  j->h->syncs--;
  napi_value undef;
  napi_get_undefined(env, &undef);
  if (status == napi_ok && j->rc == 0) {
    napi_resolve_deferred(env, j->deferred, undef);
  } else {
    char msg[512];
    snprintf(msg, sizeof msg, "avbd error %d: %s", j->rc, j->err);
    napi_value m, err;
    napi_create_string_utf8(env, msg, NAPI_AUTO_LENGTH, &m);
    napi_create_error(env, NULL, m, &err);
    napi_reject_deferred(env, j->deferred, err);
  }
  napi_delete_async_work(env, j->work);
  napi_delete_reference(env, j->ref);
  free(j);
}

static napi_value js_sync(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1] = {NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  Handle *h = get_handle(env, argv[0]);
  if (!h || !h->ctx) {
    napi_throw_type_error(env, NULL, "the first argument must be a context from create()");
    return NULL;
  }
  SyncJob *j = (SyncJob *)calloc(1, sizeof(SyncJob));
  if (!j) {
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  j->h = h;
  napi_value promise, name;
  if (napi_create_reference(env, argv[0], 1, &j->ref) != napi_ok) {
    free(j);
    napi_throw_error(env, NULL, "N-API call failed: napi_create_reference");
    return NULL;
  }
  if (napi_create_promise(env, &j->deferred, &promise) != napi_ok ||
      napi_create_string_utf8(env, "avbd_sync", NAPI_AUTO_LENGTH, &name) != napi_ok ||
      napi_create_async_work(env, NULL, name, sync_execute, sync_complete, j, &j->work) != napi_ok) {
    napi_delete_reference(env, j->ref);
    free(j);
    napi_throw_error(env, NULL, "N-API call failed: creating the sync job");
    return NULL;
  }
  if (napi_queue_async_work(env, j->work) != napi_ok) {
    napi_delete_async_work(env, j->work);
    napi_delete_reference(env, j->ref);
    free(j);
    napi_throw_error(env, NULL, "N-API call failed: napi_queue_async_work");
    return NULL;
  }
  h->syncs++;
  return promise;
}

/* Body arrays of readState / writeState: pos (n*3, required; gives n), then
 * quat (n*4), vel, angvel, prevVel (n*3 each), each optional. */
static int body_arrays(napi_env env, size_t argc, napi_value *argv, double **pos, double **quat, double **vel,
                       double **angvel, double **prev, int32_t *n) {
  const napi_typedarray_type F64 = napi_float64_array;
  void *d;
  size_t np = 0;
  if (!arr_arg(env, argc > 1 ? argv[1] : NULL, F64, "pos", -1, 0, &d, &np)) return 0;
  if (np % 3) {
    napi_throw_range_error(env, NULL, "pos must hold 3 numbers per body");
    return 0;
  }
  *pos = d;
  const long long N = (long long)(np / 3);
  if (!arr_arg(env, argc > 2 ? argv[2] : NULL, F64, "quat", 4 * N, 1, &d, NULL)) return 0;
  *quat = d;
  if (vel) {
    if (!arr_arg(env, argc > 3 ? argv[3] : NULL, F64, "vel", 3 * N, 1, &d, NULL)) return 0;
    *vel = d;
    if (!arr_arg(env, argc > 4 ? argv[4] : NULL, F64, "angvel", 3 * N, 1, &d, NULL)) return 0;
    *angvel = d;
    if (!arr_arg(env, argc > 5 ? argv[5] : NULL, F64, "prevVel", 3 * N, 1, &d, NULL)) return 0;
    *prev = d;
  }
  *n = (int32_t)N;
  return 1;
}

/* readState(ctx, pos, quat, vel, angvel, prevVel) into Float64Arrays (GpuBackend.readState, backend.js:2154-2192) */
static napi_value js_read_state(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value argv[6] = {NULL, NULL, NULL, NULL, NULL, NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  double *pos, *quat, *vel, *angvel, *prev;
  int32_t n;
  if (!body_arrays(env, argc, argv, &pos, &quat, &vel, &angvel, &prev, &n)) return NULL;
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int rc = avbd_read_state(ctx, pos, quat, vel, angvel, prev, n);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* writeState(ctx, pos, quat, vel, angvel, prevVel) */
static napi_value js_write_state(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value argv[6] = {NULL, NULL, NULL, NULL, NULL, NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  double *pos, *quat, *vel, *angvel, *prev;
  int32_t n;
  if (!body_arrays(env, argc, argv, &pos, &quat, &vel, &angvel, &prev, &n)) return NULL;
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int rc = avbd_write_state(ctx, pos, quat, vel, angvel, prev, n);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* readJoints(ctx, out Float64Array of 16 per joint or spring) and
 * readStepStart(ctx, pos, quat): the rest of syncToCPU (backend.js:1984-2039) */
static napi_value js_read_joints(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2] = {NULL, NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  void *d;
  size_t n = 0;
  if (!arr_arg(env, argc > 1 ? argv[1] : NULL, napi_float64_array, "out", -1, 0, &d, &n)) return NULL;
  if (n % 16) {
    napi_throw_range_error(env, NULL, "out must hold 16 numbers per joint");
    return NULL;
  }
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int rc = avbd_read_joints(ctx, (double *)d, (int32_t)(n / 16));
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* setJointAnchor(ctx, index, rA Float64Array[3]) (GpuBackend.updateJointAnchor) */
static napi_value js_set_joint_anchor(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3] = {NULL, NULL, NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  int32_t idx = -1;
  if (argc > 1) CHECK(env, napi_get_value_int32(env, argv[1], &idx));
  void *d;
  if (!arr_arg(env, argc > 2 ? argv[2] : NULL, napi_float64_array, "rA", 3, 0, &d, NULL)) return NULL;
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int rc = avbd_set_joint_anchor(ctx, idx, (const double *)d);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

static napi_value js_read_step_start(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3] = {NULL, NULL, NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  double *pos, *quat;
  int32_t n;
  if (!body_arrays(env, argc, argv, &pos, &quat, NULL, NULL, NULL, &n)) return NULL;
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int rc = avbd_read_step_start(ctx, pos, quat, n);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* exportPoses(ctx, alpha, out Float32Array of 8 per body): the poses a
 * renderer draws at a fraction alpha of the last step (renderer.js:1327-1383) */
static napi_value js_export_poses(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3] = {NULL, NULL, NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  double alpha = 1.0;
  if (argc > 1) CHECK(env, napi_get_value_double(env, argv[1], &alpha));
  void *d;
  size_t n = 0;
  if (!arr_arg(env, argc > 2 ? argv[2] : NULL, napi_float32_array, "out", -1, 0, &d, &n)) return NULL;
  if (n % 8) {
    napi_throw_range_error(env, NULL, "out must hold 8 numbers per body");
    return NULL;
  }
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int rc = avbd_export_poses(ctx, alpha, (float *)d, (int32_t)(n / 8), 0);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* readColoring(ctx, colorOf Int32Array) -> numColors (GpuBackend.readColoring, backend.js:2053-2151) */
static napi_value js_read_coloring(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2] = {NULL, NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  void *d;
  size_t n = 0;
  if (!arr_arg(env, argc > 1 ? argv[1] : NULL, napi_int32_array, "colorOf", -1, 0, &d, &n)) return NULL;
  int32_t k = 0;
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int rc = avbd_read_coloring(ctx, (int32_t *)d, (int32_t)n, &k);
  if (rc) return throw_avbd(env, ctx, rc);
  napi_value out;
  CHECK(env, napi_create_int32(env, k, &out));
  return out;
}

/* uploadColoring(ctx, colorOf Int32Array | null): the caller's colouring for
 * the following steps, or none (GpuBackend.uploadColoring, backend.js:1915-1931) */
static napi_value js_upload_coloring(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2] = {NULL, NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  void *d;
  size_t n = 0;
  if (!arr_arg(env, argc > 1 ? argv[1] : NULL, napi_int32_array, "colorOf", -1, 1, &d, &n)) return NULL;
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int rc = avbd_upload_coloring(ctx, (const int32_t *)d, (int32_t)n);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* setBodyOrder(ctx, rank Int32Array | null): the colouring's claim order
 * (the counterpart of Solver.bodyOrder, solver.js:225-233) */
static napi_value js_set_body_order(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2] = {NULL, NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  void *d;
  size_t n = 0;
  if (!arr_arg(env, argc > 1 ? argv[1] : NULL, napi_int32_array, "rank", -1, 1, &d, &n)) return NULL;
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int rc = avbd_set_body_order(ctx, (const int32_t *)d, (int32_t)n);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* readPairs(ctx, Int32Array|null) -> count */
static napi_value js_read_pairs(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2] = {NULL, NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  void *d;
  size_t n = 0;
  if (!arr_arg(env, argc > 1 ? argv[1] : NULL, napi_int32_array, "pairs", -1, 1, &d, &n)) return NULL;
  if (n % 2) {
    napi_throw_range_error(env, NULL, "pairs must hold 2 numbers per pair");
    return NULL;
  }
  int32_t count = 0;
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int rc = avbd_read_pairs(ctx, (int32_t *)d, (int32_t)(n / 2), &count);
  if (rc) return throw_avbd(env, ctx, rc);
  napi_value out;
  CHECK(env, napi_create_int32(env, count, &out));
  return out;
}

/* readCounters(ctx) -> object (GpuBackend.lastStats, backend.js:309-311) */
static napi_value js_read_counters(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1] = {NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  avbd_counters c;
  int rc = avbd_read_counters(ctx, &c);
  if (rc) return throw_avbd(env, ctx, rc);
  napi_value o, v;
  CHECK(env, napi_create_object(env, &o));
#define PUT(name, val)                                        \
  CHECK(env, napi_create_double(env, (double)(val), &v));     \
  CHECK(env, napi_set_named_property(env, o, name, v));
  PUT("steps", c.steps);
  PUT("bodies", c.bodies);
  PUT("dynamicBodies", c.dynamic_bodies);
  PUT("pairs", c.pairs);
  PUT("newPairs", c.new_pairs);
  PUT("manifolds", c.manifolds);
  PUT("contacts", c.contacts);
  PUT("joints", c.joints);
  PUT("colors", c.colors);
  PUT("colorRounds", c.color_rounds);
  PUT("overflow", c.overflow);
#undef PUT
  return o;
}

/* setProfiling(ctx, on); readProfile(ctx) -> Float32Array of phase ms (backend.js:1836-1891) */
static napi_value js_set_profiling(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2] = {NULL, NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  bool on = false;
  if (argc > 1) napi_get_value_bool(env, argv[1], &on);
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int rc = avbd_set_profiling(ctx, on ? 1 : 0);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

static napi_value js_read_profile(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1] = {NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  napi_value ab, arr;
  void *data;
  CHECK(env, napi_create_arraybuffer(env, sizeof(float) * AVBD_NUM_PHASES, &data, &ab));
  int32_t steps = 0;
  int rc = avbd_read_profile(ctx, (float *)data, AVBD_NUM_PHASES, &steps);
  if (rc) return throw_avbd(env, ctx, rc);
  CHECK(env, napi_create_typedarray(env, napi_float32_array, AVBD_NUM_PHASES, ab, 0, &arr));
  return arr;
}

/* Typed array of `count` elements of `type` (element size `es`) whose data is
 * returned in *data. */
static napi_value new_typed(napi_env env, napi_typedarray_type type, size_t es, size_t count, void **data) {
  napi_value ab, arr;
  if (napi_create_arraybuffer(env, es * (count ? count : 1), data, &ab) != napi_ok) return NULL;
  if (napi_create_typedarray(env, type, count, ab, 0, &arr) != napi_ok) return NULL;
  return arr;
}

/* readManifolds(ctx) -> {head Int32Array (4 per manifold: A, B, contacts,
 * flags), basisMu Float64Array (10 per manifold), contacts Float64Array (16 per
 * contact), features Uint32Array (8 per manifold), material Float64Array (3
 * per manifold: static, dynamic friction, contact distance), bias Float64Array
 * (1 per contact: the restitution target)}: every live manifold in
 * force-list order (avbd_read_manifolds, avbd_read_forces,
 * avbd_read_contact_model), for GpuBackend.readColoring's slots
 * (backend.js:2053-2151). */
static napi_value js_read_manifolds(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1] = {NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int32_t m = 0, c = 0, nf = 0;
  int rc = avbd_read_manifolds(ctx, NULL, NULL, NULL, 0, 0, &m, &c);
  if (rc) return throw_avbd(env, ctx, rc);
  rc = avbd_read_forces(ctx, NULL, NULL, 0, &nf);
  if (rc) return throw_avbd(env, ctx, rc);
  void *ph, *pb, *pc, *pk, *pf, *pm, *pr;
  napi_value head = new_typed(env, napi_int32_array, 4, (size_t)m * 4, &ph);
  napi_value bm = new_typed(env, napi_float64_array, 8, (size_t)m * 10, &pb);
  napi_value con = new_typed(env, napi_float64_array, 8, (size_t)c * 16, &pc);
  napi_value kab = new_typed(env, napi_int32_array, 4, (size_t)nf * 4, &pk);
  napi_value ft = new_typed(env, napi_uint32_array, 4, (size_t)nf * 8, &pf);
  napi_value mat = new_typed(env, napi_float64_array, 8, (size_t)m * 3, &pm);
  napi_value bias = new_typed(env, napi_float64_array, 8, (size_t)c, &pr);
  if (!head || !bm || !con || !kab || !ft || !mat || !bias) {
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  rc = avbd_read_manifolds(ctx, (int32_t *)ph, (double *)pb, (double *)pc, m, c, &m, &c);
  if (rc) return throw_avbd(env, ctx, rc);
  rc = avbd_read_forces(ctx, (int32_t *)pk, (uint32_t *)pf, nf, &nf);
  if (rc) return throw_avbd(env, ctx, rc);
  rc = avbd_read_contact_model(ctx, (double *)pm, (double *)pr, m, c);
  if (rc) return throw_avbd(env, ctx, rc);
  napi_value o;
  CHECK(env, napi_create_object(env, &o));
  CHECK(env, napi_set_named_property(env, o, "head", head));
  CHECK(env, napi_set_named_property(env, o, "basisMu", bm));
  CHECK(env, napi_set_named_property(env, o, "contacts", con));
  CHECK(env, napi_set_named_property(env, o, "forces", kab));
  CHECK(env, napi_set_named_property(env, o, "features", ft));
  CHECK(env, napi_set_named_property(env, o, "material", mat));
  CHECK(env, napi_set_named_property(env, o, "bias", bias));
  return o;
}

/* setColoringLimits(ctx, rounds, colors): GpuBackend.colorRounds / encColors
 * (0: none, the greedy fixed point) */
static napi_value js_set_coloring_limits(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3] = {NULL, NULL, NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  int32_t rounds = 0, colors = 0;
  if (argc > 1) napi_get_value_int32(env, argv[1], &rounds);
  if (argc > 2) napi_get_value_int32(env, argv[2], &colors);
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int rc = avbd_set_coloring_limits(ctx, rounds, colors);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* readInertialTarget(ctx, pos, quat): the inertial target y (bInerP / bInerQ) */
static napi_value js_read_inertial_target(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3] = {NULL, NULL, NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  double *pos, *quat;
  int32_t n;
  if (!body_arrays(env, argc, argv, &pos, &quat, NULL, NULL, NULL, &n)) return NULL;
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int rc = avbd_read_inertial_target(ctx, pos, quat, n);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* autotune(ctx) -> {team, block, trials: Float32Array(24)}: one step that
 * times the solve's team and workgroup shapes and keeps the fastest
 * (avbd_autotune, avbd_read_launch_shape; trials[t * 4 + w] for team
 * 1, 4, 8, 16, 32, 64 and workgroup 64, 128, 256, 512, 0 where not tried) */
static napi_value js_autotune(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1] = {NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int32_t team = 0, block = 0;
  int rc = avbd_autotune(ctx, &team, NULL);
  if (rc) return throw_avbd(env, ctx, rc);
  napi_value ab, arr, o, v;
  void *data = NULL;
  CHECK(env, napi_create_arraybuffer(env, 24 * sizeof(float), &data, &ab));
  rc = avbd_read_launch_shape(ctx, &team, &block, (float *)data);
  if (rc) return throw_avbd(env, ctx, rc);
  CHECK(env, napi_create_typedarray(env, napi_float32_array, 24, ab, 0, &arr));
  CHECK(env, napi_create_object(env, &o));
  CHECK(env, napi_create_int32(env, team, &v));
  CHECK(env, napi_set_named_property(env, o, "team", v));
  CHECK(env, napi_create_int32(env, block, &v));
  CHECK(env, napi_set_named_property(env, o, "block", v));
  CHECK(env, napi_set_named_property(env, o, "trials", arr));
  return o;
}

/* setLaunchShape(ctx, team, block): fix lanes per body and per workgroup
 * (avbd_set_launch_shape) */
static napi_value js_set_launch_shape(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3] = {NULL, NULL, NULL};
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  int32_t team = 0, block = 0;
  if (argc < 3 || napi_get_value_int32(env, argv[1], &team) != napi_ok ||
      napi_get_value_int32(env, argv[2], &block) != napi_ok) {
    napi_throw_type_error(env, NULL, "setLaunchShape(ctx, team, block) takes two integers");
    return NULL;
  }
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int rc = avbd_set_launch_shape(ctx, team, block);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

static napi_value js_abi_version(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value v;
  CHECK(env, napi_create_int32(env, avbd_abi_version(), &v));
  return v;
}

static napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor d[] = {
      {"abiVersion", NULL, js_abi_version, NULL, NULL, NULL, napi_enumerable, NULL},
      {"create", NULL, js_create, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setParams", NULL, js_set_params, NULL, NULL, NULL, napi_enumerable, NULL},
      {"pack", NULL, js_pack, NULL, NULL, NULL, napi_enumerable, NULL},
      {"step", NULL, js_step, NULL, NULL, NULL, napi_enumerable, NULL},
      {"sync", NULL, js_sync, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readState", NULL, js_read_state, NULL, NULL, NULL, napi_enumerable, NULL},
      {"writeState", NULL, js_write_state, NULL, NULL, NULL, napi_enumerable, NULL},
      {"exportPoses", NULL, js_export_poses, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readJoints", NULL, js_read_joints, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readStepStart", NULL, js_read_step_start, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setJointAnchor", NULL, js_set_joint_anchor, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readColoring", NULL, js_read_coloring, NULL, NULL, NULL, napi_enumerable, NULL},
      {"uploadColoring", NULL, js_upload_coloring, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setBodyOrder", NULL, js_set_body_order, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readPairs", NULL, js_read_pairs, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readCounters", NULL, js_read_counters, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setProfiling", NULL, js_set_profiling, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readProfile", NULL, js_read_profile, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readManifolds", NULL, js_read_manifolds, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setColoringLimits", NULL, js_set_coloring_limits, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readInertialTarget", NULL, js_read_inertial_target, NULL, NULL, NULL, napi_enumerable, NULL},
      {"autotune", NULL, js_autotune, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setLaunchShape", NULL, js_set_launch_shape, NULL, NULL, NULL, napi_enumerable, NULL},
  };
  if (napi_define_properties(env, exports, sizeof d / sizeof d[0], d) != napi_ok) return NULL;
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
