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

static avbd_ctx *get_ctx(napi_env env, napi_value v) {
  void *p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok) return NULL;
  return (avbd_ctx *)p;
}

static void *typed_data(napi_env env, napi_value v, size_t *len) {
  bool is = false;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is) return NULL;
  napi_typedarray_type t;
  size_t n;
  void *data;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, v, &t, &n, &data, &ab, &off) != napi_ok) return NULL;
  if (len) *len = n;
  return data;
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
  if (data) avbd_destroy((avbd_ctx *)data);
}

/* create({device, precision, team}) -> external (GpuBackend.init, backend.js:337-469) */
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
  }
  avbd_ctx *ctx = NULL;
  int rc = avbd_create(&ctx, &cfg);
  if (rc) return throw_avbd(env, NULL, rc);
  napi_value ext;
  CHECK(env, napi_create_external(env, ctx, finalize_ctx, NULL, &ext));
  return ext;
}

/* setParams(ctx, {dt, gravity, iterations, ...}) — Solver.defaultParams names (solver.js:92-236) */
static napi_value js_set_params(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
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
  int rc = avbd_set_params(ctx, &p);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* pack(ctx, arenas) with typed arrays named as avbd_scene's fields (GpuBackend.pack, backend.js:662-1301) */
static napi_value js_pack(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  napi_value o = argv[1];
  avbd_scene s;
  memset(&s, 0, sizeof s);
  size_t n = 0, nj = 0;
  napi_value v;
#define TA(field, key, lenp)                          \
  v = get_prop(env, o, key);                          \
  s.field = v ? typed_data(env, v, lenp) : NULL;
  TA(shape, "shape", &n);
  TA(size, "size", NULL);
  TA(density, "density", NULL);
  TA(friction, "friction", NULL);
  TA(pos, "pos", NULL);
  TA(quat, "quat", NULL);
  TA(vel, "vel", NULL);
  TA(angvel, "angvel", NULL);
  TA(scene_of, "sceneOf", NULL);
  TA(joint_a, "jointA", &nj);
  TA(joint_b, "jointB", NULL);
  TA(joint_ra, "jointRA", NULL);
  TA(joint_rb, "jointRB", NULL);
  TA(joint_stiff_lin, "jointStiffLin", NULL);
  TA(joint_stiff_ang, "jointStiffAng", NULL);
  TA(joint_fracture, "jointFracture", NULL);
  TA(joint_kind, "jointKind", NULL);
#undef TA
  s.num_bodies = (int32_t)n;
  s.num_joints = s.joint_a ? (int32_t)nj : 0;
  /* pack(ctx, scene, oldIndex Int32Array): avbd_repack, keeping state */
  size_t no = 0;
  const int32_t *old_index = argc > 2 ? (const int32_t *)typed_data(env, argv[2], &no) : NULL;
  if (old_index && no != n) {
    napi_throw_range_error(env, NULL, "oldIndex needs one entry per body");
    return NULL;
  }
  int rc = old_index ? avbd_repack(ctx, &s, old_index) : avbd_pack(ctx, &s);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* step(ctx, n) (GpuBackend.step, backend.js:1393-1620) */
static napi_value js_step(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  int32_t n = 1;
  if (argc > 1) napi_get_value_int32(env, argv[1], &n);
  int rc = avbd_step(ctx, n);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* sync(ctx) -> Promise (GpuBackend.idle, backend.js:1804-1818) */
typedef struct {
  avbd_ctx *ctx;
  int rc;
  napi_deferred deferred;
  napi_async_work work;
} SyncJob;

static void sync_execute(napi_env env, void *data) {
  (void)env;
  SyncJob *j = (SyncJob *)data;
  j->rc = avbd_sync(j->ctx);
}

static void sync_complete(napi_env env, napi_status status, void *data) {
  SyncJob *j = (SyncJob *)data;
  napi_value undef;
  napi_get_undefined(env, &undef);
  if (status == napi_ok && j->rc == 0) {
    napi_resolve_deferred(env, j->deferred, undef);
  } else {
    char msg[512];
    snprintf(msg, sizeof msg, "avbd error %d: %s", j->rc, avbd_last_error(j->ctx));
    napi_value m, err;
    napi_create_string_utf8(env, msg, NAPI_AUTO_LENGTH, &m);
    napi_create_error(env, NULL, m, &err);
    napi_reject_deferred(env, j->deferred, err);
  }
  napi_delete_async_work(env, j->work);
  free(j);
}

static napi_value js_sync(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  SyncJob *j = (SyncJob *)calloc(1, sizeof(SyncJob));
  j->ctx = get_ctx(env, argv[0]);
  napi_value promise, name;
  CHECK(env, napi_create_promise(env, &j->deferred, &promise));
  CHECK(env, napi_create_string_utf8(env, "avbd_sync", NAPI_AUTO_LENGTH, &name));
  CHECK(env, napi_create_async_work(env, NULL, name, sync_execute, sync_complete, j, &j->work));
  CHECK(env, napi_queue_async_work(env, j->work));
  return promise;
}

/* readState(ctx, pos, quat, vel, angvel, prevVel) into Float64Arrays (GpuBackend.readState, backend.js:2154-2192) */
static napi_value js_read_state(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value argv[6];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  size_t n = 0;
  double *pos = argc > 1 ? (double *)typed_data(env, argv[1], &n) : NULL;
  double *quat = argc > 2 ? (double *)typed_data(env, argv[2], NULL) : NULL;
  double *vel = argc > 3 ? (double *)typed_data(env, argv[3], NULL) : NULL;
  double *angvel = argc > 4 ? (double *)typed_data(env, argv[4], NULL) : NULL;
  double *prev = argc > 5 ? (double *)typed_data(env, argv[5], NULL) : NULL;
  int rc = avbd_read_state(ctx, pos, quat, vel, angvel, prev, (int32_t)(n / 3));
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* writeState(ctx, pos, quat, vel, angvel, prevVel) */
static napi_value js_write_state(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value argv[6];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  size_t n = 0;
  double *pos = argc > 1 ? (double *)typed_data(env, argv[1], &n) : NULL;
  double *quat = argc > 2 ? (double *)typed_data(env, argv[2], NULL) : NULL;
  double *vel = argc > 3 ? (double *)typed_data(env, argv[3], NULL) : NULL;
  double *angvel = argc > 4 ? (double *)typed_data(env, argv[4], NULL) : NULL;
  double *prev = argc > 5 ? (double *)typed_data(env, argv[5], NULL) : NULL;
  int rc = avbd_write_state(ctx, pos, quat, vel, angvel, prev, (int32_t)(n / 3));
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* readJoints(ctx, out Float64Array of 16 per joint or spring) and
 * readStepStart(ctx, pos, quat): the rest of syncToCPU (backend.js:1984-2039) */
static napi_value js_read_joints(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  size_t n = 0;
  double *out = argc > 1 ? (double *)typed_data(env, argv[1], &n) : NULL;
  int rc = avbd_read_joints(ctx, out, (int32_t)(n / 16));
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* setJointAnchor(ctx, index, rA Float64Array[3]) (GpuBackend.updateJointAnchor) */
static napi_value js_set_joint_anchor(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  int32_t idx = -1;
  if (argc > 1) CHECK(env, napi_get_value_int32(env, argv[1], &idx));
  size_t n = 0;
  const double *rA = argc > 2 ? (const double *)typed_data(env, argv[2], &n) : NULL;
  if (!rA || n < 3) {
    napi_throw_range_error(env, NULL, "rA needs 3 numbers");
    return NULL;
  }
  int rc = avbd_set_joint_anchor(ctx, idx, rA);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

static napi_value js_read_step_start(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  size_t n = 0;
  double *pos = argc > 1 ? (double *)typed_data(env, argv[1], &n) : NULL;
  double *quat = argc > 2 ? (double *)typed_data(env, argv[2], NULL) : NULL;
  int rc = avbd_read_step_start(ctx, pos, quat, (int32_t)(n / 3));
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* exportPoses(ctx, alpha, out Float32Array of 8 per body): the poses a
 * renderer draws at a fraction alpha of the last step (renderer.js:1327-1383) */
static napi_value js_export_poses(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  double alpha = 1.0;
  if (argc > 1) CHECK(env, napi_get_value_double(env, argv[1], &alpha));
  size_t n = 0;
  float *out = argc > 2 ? (float *)typed_data(env, argv[2], &n) : NULL;
  int rc = avbd_export_poses(ctx, alpha, out, (int32_t)(n / 8), 0);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* readColoring(ctx, colorOf Int32Array) -> numColors (GpuBackend.readColoring, backend.js:2053-2151) */
static napi_value js_read_coloring(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  size_t n = 0;
  int32_t *col = (int32_t *)typed_data(env, argv[1], &n);
  int32_t k = 0;
  int rc = avbd_read_coloring(ctx, col, (int32_t)n, &k);
  if (rc) return throw_avbd(env, ctx, rc);
  napi_value out;
  CHECK(env, napi_create_int32(env, k, &out));
  return out;
}

/* uploadColoring(ctx, colorOf Int32Array | null): the caller's colouring for
 * the following steps, or none (GpuBackend.uploadColoring, backend.js:1915-1931) */
static napi_value js_upload_coloring(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  napi_valuetype t = napi_null;
  if (argc > 1) CHECK(env, napi_typeof(env, argv[1], &t));
  size_t n = 0;
  int32_t *col = (argc > 1 && t != napi_null && t != napi_undefined) ? (int32_t *)typed_data(env, argv[1], &n) : NULL;
  int rc = avbd_upload_coloring(ctx, col, (int32_t)n);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* setBodyOrder(ctx, rank Int32Array | null): the colouring's claim order
 * (the counterpart of Solver.bodyOrder, solver.js:225-233) */
static napi_value js_set_body_order(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  napi_valuetype t = napi_null;
  if (argc > 1) CHECK(env, napi_typeof(env, argv[1], &t));
  size_t n = 0;
  int32_t *rank = (argc > 1 && t != napi_null && t != napi_undefined) ? (int32_t *)typed_data(env, argv[1], &n) : NULL;
  int rc = avbd_set_body_order(ctx, rank, (int32_t)n);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

/* readPairs(ctx, Int32Array|null) -> count */
static napi_value js_read_pairs(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  size_t n = 0;
  int32_t *buf = argc > 1 ? (int32_t *)typed_data(env, argv[1], &n) : NULL;
  int32_t count = 0;
  int rc = avbd_read_pairs(ctx, buf, (int32_t)(n / 2), &count);
  if (rc) return throw_avbd(env, ctx, rc);
  napi_value out;
  CHECK(env, napi_create_int32(env, count, &out));
  return out;
}

/* readCounters(ctx) -> object (GpuBackend.lastStats, backend.js:309-311) */
static napi_value js_read_counters(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
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
  napi_value argv[2];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  bool on = false;
  napi_get_value_bool(env, argv[1], &on);
  int rc = avbd_set_profiling(ctx, on ? 1 : 0);
  if (rc) return throw_avbd(env, ctx, rc);
  return NULL;
}

static napi_value js_read_profile(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  avbd_ctx *ctx = get_ctx(env, argv[0]);
  napi_value ab, arr;
  void *data;
  CHECK(env, napi_create_arraybuffer(env, sizeof(float) * AVBD_NUM_PHASES, &data, &ab));
  int32_t steps = 0;
  int rc = avbd_read_profile(ctx, (float *)data, AVBD_NUM_PHASES, &steps);
  if (rc) return throw_avbd(env, ctx, rc);
  CHECK(env, napi_create_typedarray(env, napi_float32_array, AVBD_NUM_PHASES, ab, 0, &arr));
  return arr;
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
  };
  if (napi_define_properties(env, exports, sizeof d / sizeof d[0], d) != napi_ok) return NULL;
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
