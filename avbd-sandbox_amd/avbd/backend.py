"""HipBackend — the Python face of the C ABI, shaped like the reference's
GpuBackend (src/physics/gpu/backend.js:243-2226): pack / step / idle /
readState / readColoring / lastStats.

A scene is a dict with ``bodies`` (each: shape 'box' | 'sphere', size (full
widths; a sphere's diameter three times), density (0 = static), friction, pos,
and optional vel, quat (xyzw), angvel, scene, and the contact material of
the GPU contact law: staticFriction, dynamicFriction, restitution,
contactOffset, restOffset, as Rigid.setMaterial takes them, rigid.js:144-214)
and ``joints`` (each: a (-1 = world anchor), b, rA, rB, kLin, kAng, fracture;
'inf' or math.inf for infinite). Parameters use the names of the reference's
Solver (solver.js:92-236); contactModel 1 selects the GPU backend's contact
law (contactOffset, restOffset, restitutionThreshold, contactPersistence).
"""
import ctypes as C
import math

import numpy as np

from . import _lib as L

PARAM_NAMES = {
    "dt": "dt", "gravity": "gravity", "betaLin": "beta_lin", "betaAng": "beta_ang", "alpha": "alpha",
    "gamma": "gamma", "penaltyMin": "penalty_min", "penaltyMax": "penalty_max",
    "collisionMargin": "collision_margin", "stickThreshold": "stick_threshold", "iterations": "iterations",
    "postStabilize": "post_stabilize", "rotatedInertia": "rotated_inertia",
    "cachedContactJacobians": "cached_contact_jacobians", "colorOrder": "color_order",
    "contactModel": "contact_model", "contactOffset": "contact_offset", "restOffset": "rest_offset",
    "restitutionThreshold": "restitution_threshold", "contactPersistence": "contact_persistence",
}
MATERIAL_KEYS = ("staticFriction", "dynamicFriction", "restitution", "contactOffset", "restOffset")


class AvbdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"avbd error {code}: {msg}")
        self.code = code


def _num(x):
    if isinstance(x, str):
        return {"inf": math.inf, "-inf": -math.inf}[x]
    return float(x)


def _dp(a):
    return a.ctypes.data_as(L.pd) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(L.pi) if a is not None else None


def _rows(bodies, key, default, width):
    """(n, width) float64 array of one per-body vector field (flattened
    through fromiter: a fraction of the time of nested lists)."""
    n = len(bodies)
    if default is None:
        it = (x for b in bodies for x in b[key])
    else:
        it = (x for b in bodies for x in b.get(key, default))
    return np.fromiter(it, np.float64, n * width).reshape(n, width)


def scene_arrays(scene):
    """Structure-of-arrays form of a scene dict (numpy, contiguous)."""
    bodies = scene["bodies"]
    n = len(bodies)
    A = {
        "shape": np.fromiter((L.SHAPE_SPHERE if b["shape"] == "sphere" else L.SHAPE_BOX for b in bodies), np.int32,
                             n),
        "size": _rows(bodies, "size", None, 3),
        "density": np.fromiter((b["density"] for b in bodies), np.float64, n),
        "friction": np.fromiter((b["friction"] for b in bodies), np.float64, n),
        "pos": _rows(bodies, "pos", None, 3),
        "quat": _rows(bodies, "quat", (0, 0, 0, 1), 4),
        "vel": _rows(bodies, "vel", (0, 0, 0), 3),
        "angvel": _rows(bodies, "angvel", (0, 0, 0), 3),
        "scene_of": np.fromiter((b.get("scene", 0) for b in bodies), np.int32, n),
    }
    # the GPU contact law's material, when any body sets one: static and
    # dynamic friction default to the body's friction, restitution to 0, the
    # offsets to the solver's (NaN)
    if any(b.get(k) is not None for b in bodies for k in MATERIAL_KEYS):
        def col(key, default):
            return np.fromiter((default(b) if b.get(key) is None else float(b[key]) for b in bodies), np.float64, n)
        A["static_friction"] = col("staticFriction", lambda b: float(b["friction"]))
        A["dynamic_friction"] = col("dynamicFriction", lambda b: float(b["friction"]))
        A["restitution"] = col("restitution", lambda b: 0.0)
        A["contact_offset"] = col("contactOffset", lambda b: math.nan)
        A["rest_offset"] = col("restOffset", lambda b: math.nan)
    # joints and springs, in creation order; a spring ({"type": "spring", a,
    # b, rA, rB, stiffness, rest (negative or absent: current length),
    # tearStrain (absent: never)}) rides the joint arrays (include/avbd.h)
    js = scene.get("joints", [])
    m = len(js)
    spring = [j.get("type") == "spring" for j in js]
    A.update({
        "joint_a": np.array([j["a"] for j in js], np.int32),
        "joint_b": np.array([j["b"] for j in js], np.int32),
        "joint_ra": np.array([j["rA"] for j in js], np.float64).reshape(m, 3),
        "joint_rb": np.array([j["rB"] for j in js], np.float64).reshape(m, 3),
        "joint_stiff_lin": np.array([_num(j["stiffness"] if sp else j["kLin"]) for j, sp in zip(js, spring)],
                                    np.float64),
        "joint_stiff_ang": np.array([_num(j.get("rest", -1.0) if sp else j["kAng"]) for j, sp in zip(js, spring)],
                                    np.float64),
        "joint_fracture": np.array([_num(j.get("tearStrain", "inf") if sp else j["fracture"])
                                    for j, sp in zip(js, spring)], np.float64),
        "joint_kind": np.array([L.JOINT_SPRING if sp else L.JOINT_BALL for sp in spring], np.int32),
    })
    return {k: np.ascontiguousarray(v) for k, v in A.items()}


class HipBackend:
    """One device context (stands in for GpuBackend; backend.js:243-335)."""

    def __init__(self, precision=L.PREC_MIXED, team=16, device=0, manifold_cap=0, pair_cap=0, graphs=0,
                 adjoint_small=0, adjoint_tape_mb=0, scene_solve=0):
        self.L = L.lib()
        self.ctx = C.c_void_p()
        cfg = L.Config(device=device, precision=precision, team=team, manifold_cap=manifold_cap,
                       pair_cap=pair_cap, graphs=graphs, adjoint_small=adjoint_small,
                       adjoint_tape_mb=adjoint_tape_mb, scene_solve=scene_solve)
        rc = self.L.avbd_create(C.byref(self.ctx), C.byref(cfg))
        if rc != 0:
            raise AvbdError(rc, "avbd_create failed (no GPU, or bad configuration)")
        self.n = self.nj = 0
        self.precision = precision
        self.params = L.Params()
        self.L.avbd_default_params(C.byref(self.params))

    def close(self):
        if self.ctx:
            self.L.avbd_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        if rc != 0:
            raise AvbdError(rc, self.L.avbd_last_error(self.ctx).decode())

    # --- topology (backend.js:662-1301)
    def set_params(self, **kw):
        for k, v in kw.items():
            name = PARAM_NAMES.get(k, k)
            if not hasattr(self.params, name):
                continue
            setattr(self.params, name, type(getattr(self.params, name))(v))
        self._ck(self.L.avbd_set_params(self.ctx, C.byref(self.params)))

    def pack(self, scene, params=None):
        if params:
            self.set_params(**{k: v for k, v in params.items() if k in PARAM_NAMES})
        else:
            self._ck(self.L.avbd_set_params(self.ctx, C.byref(self.params)))
        s = self._scene_struct(scene)
        self._ck(self.L.avbd_pack(self.ctx, C.byref(s)))
        self.n, self.nj = s.num_bodies, s.num_joints

    def repack(self, scene, old_index, params=None):
        """A topology edit that keeps state (GpuBackend.pack's identity-
        preserving repack, backend.js:662-1301): `scene` is the scene after
        bodies were erased and/or appended, `old_index[i]` the index body i had
        before (-1 for a new body). Survivors keep their device state,
        manifolds with warm start, joints and colours."""
        if params:
            self.set_params(**{k: v for k, v in params.items() if k in PARAM_NAMES})
        s = self._scene_struct(scene)
        oi = np.ascontiguousarray(np.asarray(old_index, dtype=np.int32))
        if oi.shape != (s.num_bodies,):
            raise AvbdError(L.AVBD_E_INVALID, "old_index needs one entry per body")
        self._ck(self.L.avbd_repack(self.ctx, C.byref(s), _ip(oi)))
        self.n, self.nj = s.num_bodies, s.num_joints

    def _scene_struct(self, scene):
        A = scene_arrays(scene)
        self._arrays = A  # keep alive for the call
        s = L.Scene()
        s.num_bodies = len(A["shape"])
        s.shape, s.size, s.density, s.friction = _ip(A["shape"]), _dp(A["size"]), _dp(A["density"]), _dp(A["friction"])
        s.pos, s.quat, s.vel, s.angvel = _dp(A["pos"]), _dp(A["quat"]), _dp(A["vel"]), _dp(A["angvel"])
        s.scene_of = _ip(A["scene_of"])
        s.num_joints = len(A["joint_a"])
        if s.num_joints:
            s.joint_a, s.joint_b = _ip(A["joint_a"]), _ip(A["joint_b"])
            s.joint_ra, s.joint_rb = _dp(A["joint_ra"]), _dp(A["joint_rb"])
            s.joint_stiff_lin, s.joint_stiff_ang = _dp(A["joint_stiff_lin"]), _dp(A["joint_stiff_ang"])
            s.joint_fracture = _dp(A["joint_fracture"])
            s.joint_kind = _ip(A["joint_kind"])
        if "static_friction" in A:
            s.static_friction, s.dynamic_friction = _dp(A["static_friction"]), _dp(A["dynamic_friction"])
            s.restitution, s.contact_offset = _dp(A["restitution"]), _dp(A["contact_offset"])
            s.rest_offset = _dp(A["rest_offset"])
        return s

    # --- stepping (backend.js:1393-1620, 1804-1818)
    def step(self, n=1):
        self._ck(self.L.avbd_step(self.ctx, int(n)))

    def idle(self):
        self._ck(self.L.avbd_sync(self.ctx))

    # --- state (backend.js:2154-2192)
    def read_state(self):
        n = self.n
        pos, quat = np.zeros((n, 3)), np.zeros((n, 4))
        vel, angvel, prev = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        self._ck(self.L.avbd_read_state(self.ctx, _dp(pos), _dp(quat), _dp(vel), _dp(angvel), _dp(prev), n))
        return dict(pos=pos, quat=quat, vel=vel, angvel=angvel, prev_vel=prev)

    def read_joints(self):
        """(num_joints, 16) per joint or spring, in the scene's order:
        penaltyLin[3] penaltyAng[3] lambdaLin[3] lambdaAng[3] broken, and a
        spring's penalty, C and peak strain (backend.js:2020-2034)."""
        out = np.zeros((self.nj, 16))
        self._ck(self.L.avbd_read_joints(self.ctx, _dp(out), self.nj))
        return out

    def set_joint_anchor(self, joint, rA):
        """New anchor rA of joint `joint` (a world point for a world-anchored
        joint) without a repack (GpuBackend.updateJointAnchor)."""
        a = np.ascontiguousarray(np.asarray(rA, dtype=np.float64).reshape(3))
        self._ck(self.L.avbd_set_joint_anchor(self.ctx, int(joint), _dp(a)))

    def read_step_start(self):
        """The pose x_t the last step started from: (pos, quat)."""
        pos, quat = np.zeros((self.n, 3)), np.zeros((self.n, 4))
        self._ck(self.L.avbd_read_step_start(self.ctx, _dp(pos), _dp(quat), self.n))
        return pos, quat

    def export_poses(self, alpha=1.0, out=None):
        """Per body {p, 0, q} as float32 (n, 8) at a fraction alpha of the
        last step (renderer.js:1327-1383). `out` may be a host float32 array
        of that shape, or a CUDA tensor of it (written on the device, in the
        context's stream order: call idle() before reading it elsewhere)."""
        n = self.n
        if out is not None and hasattr(out, "data_ptr") and getattr(out, "is_cuda", False):
            if tuple(out.shape) != (n, 8) or str(out.dtype) != "torch.float32" or not out.is_contiguous():
                raise AvbdError(L.AVBD_E_INVALID, "out must be a contiguous float32 (n, 8) tensor")
            self._ck(self.L.avbd_export_poses(self.ctx, float(alpha), C.c_void_p(out.data_ptr()), n, 1))
            return out
        if out is None:
            out = np.zeros((n, 8), np.float32)
        if out.dtype != np.float32 or out.shape != (n, 8) or not out.flags["C_CONTIGUOUS"]:
            raise AvbdError(L.AVBD_E_INVALID, "out must be a contiguous float32 (n, 8) array")
        self._ck(self.L.avbd_export_poses(self.ctx, float(alpha), out.ctypes.data_as(C.c_void_p), n, 0))
        return out

    def export_poses_static(self, alpha=1.0, out=None):
        """export_poses into host memory through the library's static staging
        array (avbd_export_poses_static): allocates nothing; at most
        AVBD_STATIC_POSE_BODIES bodies."""
        n = self.n
        if out is None:
            out = np.zeros((n, 8), np.float32)
        if out.dtype != np.float32 or out.shape != (n, 8) or not out.flags["C_CONTIGUOUS"]:
            raise AvbdError(L.AVBD_E_INVALID, "out must be a contiguous float32 (n, 8) array")
        self._ck(self.L.avbd_export_poses_static(self.ctx, float(alpha), out.ctypes.data_as(C.c_void_p), n))
        return out

    def pose_source(self):
        """(pose, start_pose, stride_bytes): device addresses of the current
        and step-start f64 pose records {p[3], q[4], pad}."""
        a, b, st = C.c_void_p(), C.c_void_p(), C.c_int32()
        self._ck(self.L.avbd_pose_source(self.ctx, C.byref(a), C.byref(b), C.byref(st)))
        return a.value, b.value, st.value

    def write_state(self, pos=None, quat=None, vel=None, angvel=None, prev_vel=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in
                (pos, quat, vel, angvel, prev_vel)]
        self._ck(self.L.avbd_write_state(self.ctx, *[_dp(a) for a in arrs], self.n))

    # --- verification (backend.js:2053-2151)
    def read_pairs(self):
        cnt = C.c_int32()
        self._ck(self.L.avbd_read_pairs(self.ctx, None, 0, C.byref(cnt)))
        out = np.zeros((cnt.value, 2), np.int32)
        if cnt.value:
            self._ck(self.L.avbd_read_pairs(self.ctx, _ip(out), cnt.value, C.byref(cnt)))
        return out

    def read_coloring(self):
        col = np.zeros(self.n, np.int32)
        k = C.c_int32()
        self._ck(self.L.avbd_read_coloring(self.ctx, _ip(col), self.n, C.byref(k)))
        return col, k.value

    def read_forces(self):
        cnt = C.c_int32()
        self._ck(self.L.avbd_read_forces(self.ctx, None, None, 0, C.byref(cnt)))
        kabn = np.zeros((cnt.value, 4), np.int32)
        feat = np.zeros((cnt.value, 8), np.uint32)
        if cnt.value:
            self._ck(self.L.avbd_read_forces(self.ctx, _ip(kabn), feat.ctypes.data_as(L.pu), cnt.value,
                                             C.byref(cnt)))
        return kabn, feat

    def read_manifold(self, index):
        bm = np.zeros(10)
        con = np.zeros((8, 16))
        nc = C.c_int32()
        self._ck(self.L.avbd_read_manifold(self.ctx, int(index), _dp(bm), _dp(con), C.byref(nc)))
        return dict(basis=bm[:9].reshape(3, 3), mu=bm[9], contacts=con[:nc.value])

    def upload_coloring(self, color_of):
        """Use this colouring (colour per body) on the following steps, or the
        greedy one again with None (GpuBackend.uploadColoring)."""
        if color_of is None:
            self._ck(self.L.avbd_upload_coloring(self.ctx, None, 0))
            return
        c = np.ascontiguousarray(color_of, dtype=np.int32)
        self._ck(self.L.avbd_upload_coloring(self.ctx, _ip(c), len(c)))

    def set_body_order(self, rank):
        """Claim order of the greedy colouring (rank per body), or the
        parameters' order again with None (Solver.bodyOrder's counterpart)."""
        if rank is None:
            self._ck(self.L.avbd_set_body_order(self.ctx, None, 0))
            return
        r = np.ascontiguousarray(rank, dtype=np.int32)
        self._ck(self.L.avbd_set_body_order(self.ctx, _ip(r), len(r)))

    # --- metrics / profiling (backend.js:309-324, 1836-1891)
    def set_coloring_limits(self, rounds=0, colors=0):
        """Relaxation rounds and colours per step (GpuBackend.colorRounds /
        encColors); 0 = no limit, the exact greedy colouring."""
        self._ck(self.L.avbd_set_coloring_limits(self.ctx, int(rounds), int(colors)))

    def read_inertial_target(self):
        """The inertial target y of the last step: (pos, quat)."""
        pos, quat = np.zeros((self.n, 3)), np.zeros((self.n, 4))
        self._ck(self.L.avbd_read_inertial_target(self.ctx, _dp(pos), _dp(quat), self.n))
        return pos, quat

    def read_manifolds(self):
        """Every live manifold in force-list order: head (m, 4) {A, B, contact
        count, flags}, basis_mu (m, 10), and contacts (c, 16) per contact,
        manifold after manifold: rA rB C0 penalty lambda stick."""
        m, c = C.c_int32(), C.c_int32()
        self._ck(self.L.avbd_read_manifolds(self.ctx, None, None, None, 0, 0, C.byref(m), C.byref(c)))
        head = np.zeros((m.value, 4), np.int32)
        bm = np.zeros((m.value, 10))
        con = np.zeros((c.value, 16))
        self._ck(self.L.avbd_read_manifolds(self.ctx, _ip(head), _dp(bm), _dp(con), m.value, c.value, C.byref(m),
                                            C.byref(c)))
        return head, bm, con

    def read_contact_model(self):
        """The GPU contact law's fields of the slots read_manifolds returns, in
        the same order: material (m, 3) {static friction, dynamic friction,
        contact distance} and each contact's restitution target bias (c,)."""
        m, c = C.c_int32(), C.c_int32()
        self._ck(self.L.avbd_read_manifolds(self.ctx, None, None, None, 0, 0, C.byref(m), C.byref(c)))
        mat = np.zeros((m.value, 3))
        bias = np.zeros(c.value)
        self._ck(self.L.avbd_read_contact_model(self.ctx, _dp(mat), _dp(bias), m.value, c.value))
        return mat, bias

    def counters(self):
        c = L.Counters()
        self._ck(self.L.avbd_read_counters(self.ctx, C.byref(c)))
        return {f: getattr(c, f) for f, _ in L.Counters._fields_}

    def set_profiling(self, on=True):
        self._ck(self.L.avbd_set_profiling(self.ctx, int(on)))

    def profile(self):
        ms = (C.c_float * L.NUM_PHASES)()
        steps = C.c_int32()
        self._ck(self.L.avbd_read_profile(self.ctx, ms, L.NUM_PHASES, C.byref(steps)))
        return dict(zip(L.PHASES, [float(x) for x in ms])), steps.value

    def kernel_stats(self, kernel):
        """(total ms, launches, algorithmic bytes) of a kernel while profiling."""
        ms, n, b = C.c_double(), C.c_int64(), C.c_double()
        self._ck(self.L.avbd_read_kernel_stats(self.ctx, int(kernel), C.byref(ms), C.byref(n), C.byref(b)))
        return ms.value, n.value, b.value

    def autotune(self):
        """One step that also times the primal sweep for each team shape and
        keeps the fastest; returns (team, {team: ms})."""
        t = C.c_int32()
        ms = (C.c_float * 6)()
        self._ck(self.L.avbd_autotune(self.ctx, C.byref(t), ms))
        return t.value, dict(zip((1, 4, 8, 16, 32, 64), [float(x) for x in ms]))

    def launch_shape(self):
        """(team, block, trials): lanes per body and per workgroup in use, and
        the last autotune's times {(team, block): ms} (tried pairs only)."""
        t, b = C.c_int32(), C.c_int32()
        g = (C.c_float * 24)()
        self._ck(self.L.avbd_read_launch_shape(self.ctx, C.byref(t), C.byref(b), g))
        trials = {}
        for i, tm in enumerate((1, 4, 8, 16, 32, 64)):
            for j, blk in enumerate((64, 128, 256, 512)):
                if g[i * 4 + j] > 0:
                    trials[(tm, blk)] = float(g[i * 4 + j])
        return t.value, b.value, trials

    def set_launch_shape(self, team, block):
        """Fix lanes per body and per workgroup (see launch_shape)."""
        self._ck(self.L.avbd_set_launch_shape(self.ctx, int(team), int(block)))

    # --- adjoint (no counterpart in the reference)
    def adjoint_forward(self, steps, ckpt_every=10):
        """Step `steps` times keeping checkpoints every `ckpt_every` steps."""
        self._ck(self.L.avbd_adjoint_forward(self.ctx, int(steps), int(ckpt_every)))

    def adjoint_backward(self, dL_dpos):
        """Gradients of a loss of the final positions (dL_dpos: n x 3) with
        respect to each body's friction and initial linear velocity."""
        g = np.ascontiguousarray(dL_dpos, dtype=np.float64).reshape(self.n, 3)
        mu, v0 = np.zeros(self.n), np.zeros((self.n, 3))
        self._ck(self.L.avbd_adjoint_backward(self.ctx, _dp(g), _dp(mu), _dp(v0)))
        return mu, v0
