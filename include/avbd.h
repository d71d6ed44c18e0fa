/*
 * avbd.h — C ABI of the MI355X (gfx950) AVBD rigid-body step.
 *
 * This library is a third backend beside the reference sandbox's f64 CPU
 * engine (src/physics/solver.js) and its WebGPU backend
 * (src/physics/gpu/backend.js, class GpuBackend). Each entry point below names
 * the GpuBackend member it stands in for. The Node-API addon
 * (avbd-sandbox_amd/js/) and the Python binding (avbd-sandbox_amd/avbd/) both
 * call exactly these functions.
 *
 * Conventions
 *  - Every call returns 0 on success and a negative AVBD_E_* code on failure;
 *    avbd_last_error() then holds a message. No exception crosses the ABI.
 *  - Host arrays are caller-owned and copied in or out; device memory is owned
 *    by the context. One context per host thread; calls are not re-entrant.
 *  - Units and frames follow the reference: z-up, metres, kilograms, seconds,
 *    quaternions as x, y, z, w (scalar last), 3x3 matrices row-major.
 *  - avbd_step() only enqueues work on the context's HIP stream (the
 *    reference's step() never waits either, backend.js:1606-1611); the read
 *    functions synchronise.
 */
#ifndef AVBD_H
#define AVBD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AVBD_ABI_VERSION 6

/* Error codes. */
#define AVBD_OK 0
#define AVBD_E_INVALID (-1)   /* bad argument                               */
#define AVBD_E_HIP (-2)       /* a HIP runtime call failed                  */
#define AVBD_E_CAPACITY (-3)  /* a capacity was exceeded (see counters)     */
#define AVBD_E_STATE (-4)     /* call not valid in the context's state      */

/* Precision of the step.
 *  EXACT: every quantity in f64; the contributions to each body's system are
 *         added in the reference's force order (a team of lanes evaluates a
 *         body's contacts in parallel, one lane per matrix entry adds them in
 *         order): results equal the f64 CPU engine's, bit for bit.
 *  F64:   f64 throughout, lanes of a team assemble a body's system in
 *         parallel (tree-reduced sums).
 *  MIXED: contact records and contact arithmetic in f32, body state, world
 *         positions and the 6x6 LDL^T in f64. The throughput mode. */
#define AVBD_PREC_EXACT 0
#define AVBD_PREC_F64 1
#define AVBD_PREC_MIXED 2

/* Body shapes (rigid.js:34 `shape`). */
#define AVBD_SHAPE_BOX 0
#define AVBD_SHAPE_SPHERE 1

typedef struct avbd_ctx avbd_ctx;

/* Context configuration (stands in for GpuBackend's constructor and
 * capacity plan, backend.js:243-335 and _allocate 499-659). Zero fields take
 * defaults. */
typedef struct avbd_config {
  int32_t device;          /* HIP device ordinal                            */
  int32_t precision;       /* AVBD_PREC_*                                   */
  int32_t team;            /* lanes per body in the primal solve (0: auto);
                              AVBD_PREC_EXACT: 1 assembles each body's
                              system on one lane, anything else uses the
                              team sweep (the same bits)                    */
  int32_t manifold_cap;    /* max live + new manifolds (0: 32 per body); a
                              step that overflows it fails with
                              AVBD_E_CAPACITY and the next avbd_pack or
                              avbd_repack grows it to 1.25x what was asked */
  int32_t pair_cap;        /* max broad-phase pairs (0: 32 per body); grows
                              the same way                                  */
  int32_t graphs;          /* replayed hipGraphs: 0 auto (the solve loop of
                              each step, unless profiling); AVBD_GRAPHS_STEP
                              (1) whole steps, broad phase to velocities,
                              with no host read inside a step (see
                              avbd_step); -1 never                          */
  int32_t adjoint_small;   /* largest dynamic-body count whose adjoint step
                              is reversed in one workgroup (0: 64; -1:
                              never, always one kernel per phase)           */
  int32_t adjoint_tape_mb; /* device memory in MB the adjoint forward may
                              fill with the whole trajectory's tape, so that
                              the backward needs no segment re-runs (0:
                              4096; -1: never, re-run every checkpoint
                              segment)                                      */
  int32_t scene_solve;     /* batches of scenes (scene_of): the solve of each
                              step as one launch with one workgroup per
                              scene (0: auto, when there are 2 or more scenes
                              of at most 4096 bodies each, the joints are
                              grouped by scene and post-stabilisation is off;
                              -1: never, one launch per colour and
                              iteration)                                    */
} avbd_config;

/* avbd_config.graphs: whole steps as replayed graphs. */
#define AVBD_GRAPHS_STEP 1

/* Solver parameters (Solver.defaultParams, solver.js:92-236; GPU uniform
 * block GU_FIELDS, layout.js:223-252). */
typedef struct avbd_params {
  double dt;               /* 1/60                                          */
  double gravity;          /* along z; -10 in the reference default         */
  double beta_lin;         /* 1e4                                           */
  double beta_ang;         /* 100                                           */
  double alpha;            /* 0.95                                          */
  double gamma;            /* 0.999                                         */
  double penalty_min;      /* 1                                             */
  double penalty_max;      /* 1e10                                          */
  double collision_margin; /* 0.01                                          */
  double stick_threshold;  /* 1e-5                                          */
  int32_t iterations;      /* 10                                            */
  int32_t post_stabilize;  /* 0                                             */
  int32_t rotated_inertia; /* 1                                             */
  int32_t cached_contact_jacobians; /* 1 (the only mode the GPU implements) */
  /* Order in which bodies claim colours in the greedy colouring that
   * schedules the primal sweep. AVBD_ORDER_INDEX: body index, the CPU
   * engine's Coloring.build (coloring.js:75-92). AVBD_ORDER_PRIORITY: the
   * lowbias32 claim order of the reference GPU backend (layout.js:50-66,
   * shaders.js:1571-1585), whose dependency chains are short. Either way the
   * colouring is the exact greedy fixed point. */
  int32_t color_order;
  /* Contact law. AVBD_CONTACTS_CPU: the CPU engine's (manifold.js:214-491),
   * bit for bit in AVBD_PREC_EXACT. AVBD_CONTACTS_GPU: the reference GPU
   * backend's (shaders.js:1105-1367, 1777-1850, 2229-2264): per-body static
   * and dynamic friction with a stick-selected cone, a one-step restitution
   * target for new impacts, speculative contacts inside the pair's contact
   * skin (the sum of the bodies' contact offsets) with C0 measured against the
   * rest distance (the sum of their rest offsets) instead of the collision
   * margin, and warm starts matched geometrically when feature keys change
   * (the persistence fallback). The fields below configure it. */
  int32_t contact_model;
  double contact_offset;        /* 0.005: a body's skin unless it sets its own (solver.js:151) */
  double rest_offset;           /* -0.005: a body's rest offset unless it sets its own (solver.js:152) */
  double restitution_threshold; /* 1.0 m/s: slower impacts set no restitution target (solver.js:155) */
  double contact_persistence;   /* 0.002 m: least warm-start matching distance (solver.js:162) */
} avbd_params;

#define AVBD_CONTACTS_CPU 0
#define AVBD_CONTACTS_GPU 1

#define AVBD_ORDER_INDEX 0
#define AVBD_ORDER_PRIORITY 1
#define AVBD_ORDER_CUSTOM   2   /* a caller's order (avbd_set_body_order) */

/* Fills *p with the reference defaults (solver.js:92-236). */
void avbd_default_params(avbd_params *p);

/* Scene, structure of arrays (stands in for the arenas GpuBackend.pack builds
 * from the Solver scene graph, backend.js:662-1301; the record layout of
 * test/parity/reference_main.cpp:45-130). Bodies are numbered in creation
 * order; the step's Gauss-Seidel schedule depends on it. */
typedef struct avbd_scene {
  int32_t num_bodies;
  const int32_t *shape;       /* [n] AVBD_SHAPE_*                            */
  const double *size;         /* [n*3] full widths; spheres: diameter x3     */
  const double *density;      /* [n] 0 = static                             */
  const double *friction;     /* [n]                                        */
  const double *pos;          /* [n*3]                                      */
  const double *quat;         /* [n*4] xyzw, or NULL for identity           */
  const double *vel;          /* [n*3] or NULL                              */
  const double *angvel;       /* [n*3] or NULL                              */
  const int32_t *scene_of;    /* [n] batch scene id (bodies of different
                                 scenes never interact), or NULL: one scene */
  int32_t num_joints;         /* joint.js:78-113                            */
  const int32_t *joint_a;     /* [J] body index, -1 = world anchor          */
  const int32_t *joint_b;     /* [J]                                        */
  const double *joint_ra;     /* [J*3] anchor in A (world point if a = -1)  */
  const double *joint_rb;     /* [J*3] anchor in B                          */
  const double *joint_stiff_lin; /* [J] INFINITY = hard                     */
  const double *joint_stiff_ang; /* [J] 0 = no angular rows                 */
  const double *joint_fracture;  /* [J] INFINITY = unbreakable              */
  /* [J] or NULL (all ball joints): AVBD_JOINT_BALL, or AVBD_JOINT_SPRING for a
   * distance spring (spring.js:63-105), whose entries read joint_a/joint_b
   * (both bodies), joint_ra/joint_rb (body-frame anchors), joint_stiff_lin as
   * the stiffness, joint_stiff_ang as the rest length (negative: the anchors'
   * current distance) and joint_fracture as the tensile strain at which it
   * tears (INFINITY: never). Joints and springs are forces in the order given,
   * which is their creation order. */
  const int32_t *joint_kind;
  /* The contact material of AVBD_CONTACTS_GPU (Rigid.setMaterial,
   * rigid.js:144-214), each [n] or NULL: static and dynamic friction (NULL:
   * the body's friction), restitution in [0, 1] (NULL: 0), the contact offset
   * and the rest offset (NULL or NaN: the solver's contact_offset /
   * rest_offset). As setMaterial, pack rejects a negative or non-finite
   * coefficient, dynamic above static friction, a negative contact offset and
   * a rest offset above the body's own contact offset. A rest offset resolved
   * above the resolved contact offset is lowered to it, and a contact offset
   * is at least 0 (bodyContactSignature, backend.js:156-185). The CPU law
   * reads friction alone. */
  const double *static_friction;
  const double *dynamic_friction;
  const double *restitution;
  const double *contact_offset;
  const double *rest_offset;
} avbd_scene;

#define AVBD_JOINT_BALL 0
#define AVBD_JOINT_SPRING 1

/* Step counters (layout.js:171-178 CTR_*; backend.js:1622-1792 lastStats). */
typedef struct avbd_counters {
  int64_t steps;
  int32_t bodies, dynamic_bodies;
  int32_t pairs;           /* broad-phase pairs of the last step           */
  int32_t new_pairs;       /* of which created a manifold                   */
  int32_t manifolds;       /* live manifolds after force initialisation     */
  int32_t contacts;        /* contacts in them                              */
  int32_t joints;          /* live joints                                   */
  int32_t colors;          /* colours of the last step                      */
  int32_t color_rounds;    /* the colouring's rounds: relaxation rounds it
                              took, or for the one-launch colouring its
                              longest chain of claims-before neighbours     */
  int32_t overflow;        /* bit 1 pairs, 2 manifolds (layout.js:174), 4
                              colours, 8 work items, 16 an uploaded
                              colouring was not proper for a step, 32 the
                              colouring limits left it improper            */
} avbd_counters;

/* --- lifecycle: GpuBackend.init / initDevice / initPipelines (backend.js:337-469) */
int avbd_create(avbd_ctx **out, const avbd_config *cfg);
int avbd_destroy(avbd_ctx *ctx);
const char *avbd_last_error(const avbd_ctx *ctx);
int avbd_abi_version(void);

/* --- topology: GpuBackend.pack / flushTopology (backend.js:662-1301, 1898-1915).
 * Replaces the whole scene; warm-start state starts empty, as for a fresh
 * Solver. */
int avbd_set_params(avbd_ctx *ctx, const avbd_params *p);
int avbd_pack(avbd_ctx *ctx, const avbd_scene *scene);

/* Topology edit that keeps state (GpuBackend.pack's identity-preserving
 * repack, backend.js:662-1301): the scene after bodies were erased
 * (Solver.removeBodies / Rigid.destroy, solver.js:261-300, rigid.js:270-274)
 * and new ones appended (new Rigid(...)). old_index[i] is the index body i had
 * in the previous scene, or -1 for a new body; survivors keep their relative
 * order (AVBD_E_INVALID otherwise). Survivors keep their device state (pose,
 * step-start pose, velocities; the description's pos/quat/vel/angvel for them
 * are not read), their contact manifolds with every warm-start field (feature
 * keys, anchors, penalties, multipliers, stick flags) and their colours;
 * manifolds and joints touching an erased body are dropped, keeping the order
 * of the rest. The description's joints must be exactly the surviving joints
 * in order (their state, a fracture included, is kept): a joint added to a
 * running scene follows its contacts in the reference's force list, which the
 * packed layout (joints first) cannot express, so joints are added with
 * avbd_pack. Properties (size, density, friction, shape) come from the
 * description. On an error after validation the context holds no scene. */
int avbd_repack(avbd_ctx *ctx, const avbd_scene *scene, const int32_t *old_index);

/* --- stepping: GpuBackend.step (backend.js:1393-1620); avbd_sync is idle()
 * (backend.js:1804-1818).
 * With avbd_config.graphs = AVBD_GRAPHS_STEP, every step after the first
 * since a pack (which reads its counts on the host and sets the launch
 * shapes) is one replayed graph that takes every count from device memory,
 * as the reference's step takes its dispatch sizes from indirect buffers
 * (backend.js:1496-1535): no host read, at most 4 steps in flight. The
 * colouring then runs a fixed number of rounds and the sweep a fixed number
 * of colours per iteration: the caller's limits (avbd_set_coloring_limits,
 * GpuBackend.colorRounds / encColors) or, without them, one colour and four
 * rounds above what the last step read needed. A step that needed more
 * reports it (bit 32 of avbd_counters.overflow, as a limited colouring does)
 * and the following steps launch more. A step that exceeds the pair or
 * manifold capacity drops the excess (bits 1, 2 of the overflow word, as the
 * reference's steps do); the next avbd_step returns AVBD_E_CAPACITY and the
 * next avbd_pack or avbd_repack grows the capacity. Every other call first
 * waits for the steps in flight. Profiling, recording for the adjoint, an
 * uploaded colouring and the autotuner use steps with host reads. */
int avbd_step(avbd_ctx *ctx, int32_t nsteps);
int avbd_sync(avbd_ctx *ctx);

/* --- state: GpuBackend.readState / syncToCPU (backend.js:2154-2192, 1984-2039).
 * Any pointer may be NULL. n must equal the packed body count. */
int avbd_read_state(avbd_ctx *ctx, double *pos, double *quat, double *vel, double *angvel,
                    double *prev_vel, int32_t n);
/* --- poses for drawing. A renderer draws each body at a fraction alpha of
 * the last step, blended from the pose x_t at the step's start: positions
 * lerped, orientations nlerped along the shorter arc and renormalised, in f64
 * and stored as f32 (renderer.js:1327-1383; the WebGPU renderer reads x_t from
 * the body arena's bInitP/bInitQ, boot.js:2223-2231). avbd_export_poses writes
 * that blend, per body 8 floats {p[3], 0, q[4]}; alpha is clamped to [0, 1]
 * and from 1 on (or NaN) the current pose is copied. dst is host memory, or,
 * with dst_on_device nonzero, 16-byte aligned device memory on the context's
 * device, written in stream order (avbd_sync waits for it). n must equal the
 * packed body count. */
int avbd_export_poses(avbd_ctx *ctx, double alpha, float *dst, int32_t n, int32_t dst_on_device);
/* The same poses into host memory without allocating anything (no
 * counterpart in the reference): for callers that cannot allocate device
 * memory, such as a render or serving loop that forbids it. The poses are
 * staged in a static device array that each GPU holds for the life of the
 * process; n (the packed body count) must not exceed
 * AVBD_STATIC_POSE_BODIES (AVBD_E_INVALID). Calls on one GPU share that
 * array: calls that overlap on one GPU, from several threads or contexts,
 * race. */
#define AVBD_STATIC_POSE_BODIES 65536
int avbd_export_poses_static(avbd_ctx *ctx, double alpha, float *dst, int32_t n);
/* The body arena for a renderer on the same device that reads poses in place
 * (setPoseSourceExternal(buffer, posBase, quatBase, prevPosBase, prevQuatBase),
 * renderer_webgpu.js:1627): the device addresses of the current pose and of
 * the step-start pose x_t, each a record of f64 {p[3], q[4], pad} every
 * *stride_bytes. Valid until the next avbd_pack or avbd_destroy; read them
 * after avbd_sync. Any output pointer may be NULL. */
int avbd_pose_source(avbd_ctx *ctx, const double **pose, const double **start_pose, int32_t *stride_bytes);

/* The rest of syncToCPU (backend.js:1984-2039): the device-owned constraint
 * state and the step-start poses, for a host that mirrors them.
 * avbd_read_joints: per entry of the scene's joint arrays (n = num_joints), 16
 * doubles: penaltyLin[3] penaltyAng[3] lambdaLin[3] lambdaAng[3] (joints),
 * broken (1 once fractured or torn), then for a spring its penalty, last C
 * and peak tensile strain. avbd_read_step_start: the pose x_t the last step
 * started from (initialLin / initialAng); either pointer may be NULL. */
int avbd_read_joints(avbd_ctx *ctx, double *out, int32_t n);
/* GpuBackend.updateJointAnchor (backend.js:1969-1977): a new anchor rA for
 * joint `joint` (its index in the scene's joint arrays; the world point of a
 * world-anchored joint, as a mouse drag moves it) without a repack. Ordered
 * before the next step. */
int avbd_set_joint_anchor(avbd_ctx *ctx, int32_t joint, const double *rA);
int avbd_read_step_start(avbd_ctx *ctx, double *pos, double *quat, int32_t n);

/* Overwrites body state (the device-authoritative half of syncToCPU's
 * inverse, backend.js:662-736). */
int avbd_write_state(avbd_ctx *ctx, const double *pos, const double *quat, const double *vel,
                     const double *angvel, const double *prev_vel, int32_t n);

/* --- verification: GpuBackend.readColoring (backend.js:2053-2151).
 * Broad-phase pairs of the last step as (A, B) index pairs, A > B, in the
 * order A descending then B descending (broadphase.js:236-243). */
int avbd_read_pairs(avbd_ctx *ctx, int32_t *pairs, int32_t cap, int32_t *count);
/* Colouring used by the last step: colour per body (-1 static), and the
 * colour count (coloring.js:43-118). */
int avbd_read_coloring(avbd_ctx *ctx, int32_t *color_of, int32_t n, int32_t *num_colors);
/* Live force list in the order the step holds it (solver.js forces[]):
 * per force kind (0 manifold, 1 joint, 2 spring), body A (-1 world), body B and contact
 * count; `features` receives 8 words per force (contact feature keys,
 * collide.js:254-285). Either output may be NULL. */
int avbd_read_forces(avbd_ctx *ctx, int32_t *kind_a_b_nc, uint32_t *features, int32_t cap,
                     int32_t *count);
/* Contact records of force `index` of that list: per contact 16 doubles
 * rA[3] rB[3] C0[3] penalty[3] lambda[3] stick (manifold.js:34-70), and the
 * basis (9 doubles, row 0 the normal, B to A) and friction coefficient. */
int avbd_read_manifold(avbd_ctx *ctx, int32_t index, double *basis_mu, double *contacts,
                       int32_t *num_contacts);

/* Every live contact manifold at once, in force-list order (the manifolds of
 * avbd_read_forces' list, after its joints), the slots GpuBackend.readColoring
 * returns as slotInfo (backend.js:2053-2151): per manifold `head` {body A, body
 * B, contact count, flags (1: a sphere is involved)} and `basis_mu` (the basis,
 * row 0 the normal B to A, then the friction coefficient; 10 doubles); per
 * contact, manifold after manifold, `contacts` 16 doubles as in
 * avbd_read_manifold. At most cap manifolds and contact_cap contacts are
 * written; *count and *num_contacts receive the live totals. Any output may
 * be NULL. */
int avbd_read_manifolds(avbd_ctx *ctx, int32_t *head, double *basis_mu, double *contacts, int32_t cap,
                        int32_t contact_cap, int32_t *count, int32_t *num_contacts);

/* The fields AVBD_CONTACTS_GPU adds to the slots of avbd_read_manifolds,
 * in the same order (readColoring's slotInfo staticFriction, dynamicFriction
 * and each contact's restitutionBias, backend.js:2053-2151): per manifold
 * `material` {static friction, dynamic friction, contact distance} (3
 * doubles), per contact its restitution target `bias`. Under
 * AVBD_CONTACTS_CPU they read {friction, 0, 0} and 0. */
int avbd_read_contact_model(avbd_ctx *ctx, double *material, double *bias, int32_t cap, int32_t contact_cap);

/* --- metrics: GpuBackend.lastStats (backend.js:309-311, 1622-1792) */
int avbd_read_counters(avbd_ctx *ctx, avbd_counters *out);

/* --- profiling: GpuBackend.setProfiling / lastProfile (backend.js:1836-1891).
 * When on, each step records HIP-event times of its phases; ms receives
 * AVBD_NUM_PHASES averages over the steps since profiling was turned on. */
#define AVBD_NUM_PHASES 6 /* broad, narrow, topology, colour, solve, total */
int avbd_set_profiling(avbd_ctx *ctx, int32_t on);
int avbd_read_profile(avbd_ctx *ctx, float *ms, int32_t n, int32_t *steps);

/* Per-kernel timing while profiling is on (HIP events around each launch of
 * the kernel on the context's stream), with the algorithmic bytes each launch
 * moves: the unique body, manifold and contact records it must read or write
 * under the library's data layout (per-unit figures documented with the
 * layout). No counterpart in the reference beyond lastProfile. */
#define AVBD_K_PRIMAL 0
#define AVBD_K_DUAL 1
#define AVBD_K_NARROW 2
#define AVBD_K_POSES 3   /* avbd_export_poses */
#define AVBD_NUM_KERNELS 4
int avbd_read_kernel_stats(avbd_ctx *ctx, int32_t kernel, double *ms_total, int64_t *launches,
                           double *bytes_total);

/* Autotuner (no counterpart in the reference). Runs one step during which
 * the primal sweep of the first iteration is timed for every team shape
 * (lanes per body: 1, 4, 8, 16, 32, 64) on the step's own data, restoring the
 * body state between trials; the step then completes with the fastest shape,
 * which the context keeps. `team` receives it and `ms` (may be NULL) the six
 * sweep times in that order. A batch solved in one launch per step
 * (avbd_config.scene_solve) has the whole solve timed instead for lanes per
 * body 4, 8 and 16 (a workgroup of 256 lanes then updates 64, 32 or 16 bodies
 * of a colour at once); the other entries of `ms` are 0. Not available in
 * AVBD_PREC_EXACT, whose assembly order is fixed. */
int avbd_autotune(avbd_ctx *ctx, int32_t *team, float *ms);

/* Launch shape (no counterpart in the reference). avbd_autotune also times
 * each team shape at several workgroup sizes: the per-colour sweep at 64, 128
 * and 256 lanes per workgroup, the one-launch batch solve at 128, 256 and 512
 * lanes per scene (with TEAM lanes per body a scene's workgroup updates
 * block / TEAM bodies of a colour at once), and keeps the fastest pair; `ms`
 * of avbd_autotune then holds each team shape's best time.
 * avbd_read_launch_shape returns the pair in use (`team`: lanes per body,
 * `block`: lanes per workgroup of the solve's launches) and, in `grid` (may
 * be NULL; 24 floats), every trial of the last autotune as
 * [team 1, 4, 8, 16, 32, 64][workgroup 64, 128, 256, 512] in ms, 0 where not
 * tried. avbd_set_launch_shape fixes the pair (team in {1, 4, 8, 16, 32, 64},
 * block in {64, 128, 256} for the per-colour sweep; for a batch solved in one
 * launch, team in {4, 8, 16} and block in {128, 256, 512}); results do not
 * depend on the workgroup. AVBD_PREC_EXACT keeps its team (32 lanes per
 * body, 1 with avbd_config.team = 1) and takes block 64, 128 or 256 for its
 * team sweep (avbd_set_launch_shape(ctx, 32, block)). */
int avbd_read_launch_shape(avbd_ctx *ctx, int32_t *team, int32_t *block, float *grid);
int avbd_set_launch_shape(avbd_ctx *ctx, int32_t team, int32_t block);

/* Verification hook: the claim order of the greedy colouring, and so the
 * sweep's schedule (the reference CPU engine's Solver.bodyOrder,
 * solver.js:225-233 and 419-422, reproduces a colour schedule from the
 * flattened colouring; here the colouring is built in the given order).
 * rank[i] is body i's position (lower claims first, ties to the lower index);
 * NULL returns to avbd_params.color_order. Kept until changed or repacked. */
int avbd_set_body_order(avbd_ctx *ctx, const int32_t *rank, int32_t n);

/* Verification hook (GpuBackend.uploadColoring / overrideColoring,
 * backend.js:1915-1931, 299, 1508-1516): use the caller's colouring (colour of
 * each body; ignored for static bodies) instead of the greedy one on the
 * following steps, until called with NULL or repacked. Each step checks it
 * against that step's forces; a step where two bodies that share a force have
 * one colour colours greedily instead and sets bit 16 of
 * avbd_counters.overflow. AVBD_E_INVALID for a dynamic body's colour outside
 * [0, 256). */
int avbd_upload_coloring(avbd_ctx *ctx, const int32_t *color_of, int32_t n);

/* Limits on the colouring (GpuBackend.colorRounds and encColors,
 * backend.js:256-270, 1403-1430): at most `rounds` relaxation rounds per step
 * and at most `colors` colours (bodies above take the last one). 0 = no limit,
 * the default: the colouring is then the exact greedy fixed point. A limited
 * colouring may give two bodies of one force the same colour; they are then
 * updated in the same sweep, the conflict the reference's GPU tolerates
 * (layout.js:26-43), and bit 32 of avbd_counters.overflow is set. */
int avbd_set_coloring_limits(avbd_ctx *ctx, int32_t rounds, int32_t colors);

/* The inertial target y of the last step (the reference GPU's bInerP /
 * bInerQ, layout.js BODY_SECTIONS; solver.js:366-401). */
int avbd_read_inertial_target(avbd_ctx *ctx, double *pos, double *quat, int32_t n);

/* Adjoint of the step (no counterpart in the reference; the build plan's
 * adjoint boundary). avbd_adjoint_forward steps the scene nsteps times, as
 * avbd_step does, keeping a checkpoint of the carried state every ckpt_every
 * steps. avbd_adjoint_backward then returns the gradient of a loss L(final
 * positions), given dL_dpos [n*3], with respect to each body's friction
 * coefficient (dL_dmu [n]) and initial linear velocity (dL_dv0 [n*3]; pack
 * sets the velocity and the previous velocity to it). It re-runs each segment
 * from its checkpoint recording every iteration and reverses it; the context
 * ends at the forward's final state. The derivative is that of the computed
 * trajectory (a fixed iteration count), with the contact geometry (normals,
 * features, anchors) held piecewise constant and only the normal row of C0
 * differentiated for contacts whose anchors were re-derived that step.
 * Requires AVBD_PREC_EXACT or AVBD_PREC_F64, box contacts, no joints, no
 * post-stabilisation and AVBD_CONTACTS_CPU (AVBD_E_STATE otherwise). The forward must start from
 * the scene as avbd_pack set it (no step, avbd_write_state or avbd_repack
 * since; AVBD_E_STATE otherwise). A call that changes what the forward
 * stepped through or steps further (avbd_set_params, avbd_step,
 * avbd_write_state, avbd_set_joint_anchor, avbd_set_body_order,
 * avbd_upload_coloring, avbd_autotune, a pack or repack) drops the
 * recording: the backward then returns AVBD_E_STATE before any launch. The
 * tape is kept in device memory while what has been recorded fits
 * avbd_config.adjoint_tape_mb; past it the tape is released and the backward
 * re-runs each segment from its checkpoint. */
int avbd_adjoint_forward(avbd_ctx *ctx, int32_t nsteps, int32_t ckpt_every);
int avbd_adjoint_backward(avbd_ctx *ctx, const double *dL_dpos, double *dL_dmu, double *dL_dv0);

#ifdef __cplusplus
}
#endif
#endif /* AVBD_H */
