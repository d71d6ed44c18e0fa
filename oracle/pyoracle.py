"""ctypes wrapper over liboracle.so — the f64 CPU oracle.

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker or the timed CPU
reference, never as the product path.

A scene is given as a dict with ``bodies`` (shape, size, density, friction,
pos, vel, quat, angvel) and ``joints`` (a, b, rA, rB, kLin, kAng, fracture; or,
for a spring, type="spring", a, b, rA, rB, stiffness, rest, tearStrain), the
same description the HIP backend's ``pack`` takes.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# AVBD_ORACLE_LIB selects another build of the same library (the sanitized
# one of oracle/run_sanitized.sh)
LIB_PATH = os.environ.get("AVBD_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

SCHED_REVERSE, SCHED_ORDER, SCHED_COLOR = 0, 1, 2

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        d, i, v = C.c_double, C.c_int, C.c_void_p
        pd, pi = P(d), P(i)
        sig = {
            "orc_create": (v, []), "orc_destroy": (None, [v]),
            "orc_set_params": (None, [v, pd, pi]),
            "orc_add_box": (i, [v, pd, d, d, pd, pd, pd, pd]),
            "orc_add_sphere": (i, [v, d, d, d, pd, pd, pd, pd]),
            "orc_add_joint": (i, [v, i, i, pd, pd, d, d, d]),
            "orc_set_schedule": (None, [v, i, pi, i]),
            "orc_set_broadphase": (None, [v, i]),
            "orc_set_color_order": (None, [v, i]),
            "orc_color_priority": (C.c_uint, [C.c_uint]),
            "orc_set_body": (None, [v, i, pd, pd, pd, pd]),
            "orc_record_pairs": (None, [v, i]),
            "orc_step": (d, [v, i]),
            "orc_num_bodies": (i, [v]), "orc_num_dynamic": (i, [v]),
            "orc_get_state": (None, [v, pd]), "orc_get_mass": (None, [v, pd]),
            "orc_num_forces": (i, [v]),
            "orc_get_forces": (None, [v, pi, P(C.c_longlong)]),
            "orc_get_manifold": (i, [v, i, pd, P(C.c_uint), pd]),
            "orc_get_joint": (i, [v, i, pd]),
            "orc_coloring": (i, [v, pi, pi, pi]),
            "orc_last_coloring": (i, [v, pi]),
            "orc_num_pairs": (i, [v]), "orc_get_pairs": (None, [v, pi]),
            "orc_collide": (i, [v, i, i, pd, pd]),
            "orc_solve6": (None, [pd, pd, pd, pd, pd, pd, pd]),
            "orc_load_scene": (i, [v, C.c_char_p]),
            "orc_remove_bodies": (i, [v, pi, i]),
            "orc_add_spring": (i, [v, i, i, pd, pd, d, d, d]),
            "orc_get_spring": (i, [v, i, pd]),
            "orc_set_joint_anchor": (i, [v, i, pd]),
            "orc_format_state": (C.c_long, [v, i, C.c_char_p, C.c_long]),
            "orc_set_contact_model": (None, [v, i, d, d, d, d]),
            "orc_set_material": (i, [v, i, d, d, d, d, d]),
            "orc_get_manifold_gpu": (i, [v, i, pd, pd]),
            "orc_geometric_matches": (C.c_longlong, [v]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _vec(x, n):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(n))


def _num(x):
    if isinstance(x, str):
        return {"inf": math.inf, "-inf": -math.inf}[x]
    return float(x)


DEFAULT_PARAMS = dict(dt=1.0 / 60.0, gravity=-10.0, iterations=10, betaLin=1e4, betaAng=100.0,
                      alpha=0.95, gamma=0.999, penaltyMin=1.0, penaltyMax=1e10,
                      collisionMargin=0.01, stickThreshold=1e-5, postStabilize=False,
                      rotatedInertia=True, cachedContactJacobians=True)
# the reference GPU backend's contact model (solver.js:142-162); contactModel
# 0 is the CPU engine's, 1 the GPU backend's
CONTACT_PARAMS = dict(contactModel=0, contactOffset=0.005, restOffset=-0.005, restitutionThreshold=1.0,
                      contactPersistence=0.002)


def _material(b):
    """The body's GPU contact material (Rigid.setMaterial, rigid.js:144-214),
    or None when the body sets none: static and dynamic friction default to
    its friction (the constructor's, rigid.js:56-58), restitution to 0, the
    offsets (None) to the solver's."""
    keys = ("staticFriction", "dynamicFriction", "restitution", "contactOffset", "restOffset")
    if not any(b.get(k) is not None for k in keys):
        return None
    mu = float(b["friction"])
    sf = float(b["staticFriction"]) if b.get("staticFriction") is not None else mu
    df = float(b["dynamicFriction"]) if b.get("dynamicFriction") is not None else mu
    off = [math.nan if b.get(k) is None else float(b[k]) for k in ("contactOffset", "restOffset")]
    return sf, df, float(b.get("restitution") or 0.0), off[0], off[1]


class Oracle:
    """One f64 world. Mirrors the reference ``Solver`` (solver.js:59-236)."""

    def __init__(self, scene, params=None, schedule=SCHED_COLOR, broadphase="grid", color_order=0):
        self.L = lib()
        self.h = self.L.orc_create()
        p = dict(DEFAULT_PARAMS)
        p.update({k: v for k, v in (params or {}).items() if k in DEFAULT_PARAMS})
        self.params = p
        self._params_to_lib()
        cm = dict(CONTACT_PARAMS)
        cm.update({k: v for k, v in (params or {}).items() if k in CONTACT_PARAMS})
        self.L.orc_set_contact_model(self.h, int(cm["contactModel"]), float(cm["contactOffset"]),
                                     float(cm["restOffset"]), float(cm["restitutionThreshold"]),
                                     float(cm["contactPersistence"]))
        if isinstance(scene, (str, os.PathLike)):
            # a .scene file, read by the oracle's C restatement of the
            # reference loader (test/parity/reference_main.cpp:45-124)
            r = self.L.orc_load_scene(self.h, os.fsencode(scene))
            if r < 0:
                raise ValueError(f"orc_load_scene({scene}) failed: {r}")
            scene = {"bodies": [], "joints": []}
        for b in scene["bodies"]:
            q = _vec(b.get("quat", (0, 0, 0, 1)), 4)
            w = _vec(b.get("angvel", (0, 0, 0)), 3)
            vel = _vec(b.get("vel", (0, 0, 0)), 3)
            pos = _vec(b["pos"], 3)
            if b["shape"] == "sphere":
                self.L.orc_add_sphere(self.h, float(b["size"][0]), float(b["density"]), float(b["friction"]),
                                      _dp(pos), _dp(vel), _dp(q), _dp(w))
            else:
                size = _vec(b["size"], 3)
                self.L.orc_add_box(self.h, _dp(size), float(b["density"]), float(b["friction"]),
                                   _dp(pos), _dp(vel), _dp(q), _dp(w))
        for i, b in enumerate(scene["bodies"]):
            self._set_material(i, b)
        for j in scene.get("joints", []):
            rA, rB = _vec(j["rA"], 3), _vec(j["rB"], 3)
            if j.get("type") == "spring":
                r = self.L.orc_add_spring(self.h, int(j["a"]), int(j["b"]), _dp(rA), _dp(rB),
                                          _num(j["stiffness"]), _num(j.get("rest", -1.0)),
                                          _num(j.get("tearStrain", "inf")))
                if r < 0:
                    raise ValueError(f"bad spring {j}")
                continue
            r = self.L.orc_add_joint(self.h, int(j["a"]), int(j["b"]), _dp(rA), _dp(rB), _num(j["kLin"]),
                                     _num(j["kAng"]), _num(j["fracture"]))
            if r < 0:
                raise ValueError(f"bad joint {j}")
        self.set_schedule(schedule)
        self.L.orc_set_broadphase(self.h, 1 if broadphase == "brute" else 0)
        self.L.orc_set_color_order(self.h, int(color_order))

    def _params_to_lib(self):
        p = self.params
        dv = np.array([p["dt"], p["gravity"], p["betaLin"], p["betaAng"], p["alpha"], p["gamma"],
                       p["penaltyMin"], p["penaltyMax"], p["collisionMargin"], p["stickThreshold"]],
                      dtype=np.float64)
        iv = np.array([p["iterations"], int(p["postStabilize"]), int(p["rotatedInertia"]),
                       int(p["cachedContactJacobians"])], dtype=np.int32)
        self.L.orc_set_params(self.h, _dp(dv), _ip(iv))

    def set_params(self, **kw):
        """Changes solver parameters between steps (solver.iterations = 1 and
        the like)."""
        unknown = set(kw) - set(DEFAULT_PARAMS)
        if unknown:
            raise KeyError(sorted(unknown))
        self.params.update(kw)
        self._params_to_lib()

    def geometric_matches(self):
        """GPU contact model: contacts warm-started by the geometric
        persistence match (another feature key), over all steps so far."""
        return int(self.L.orc_geometric_matches(self.h))

    def _set_material(self, i, b):
        m = _material(b)
        if m is not None and self.L.orc_set_material(self.h, int(i), *m) < 0:
            raise ValueError(f"body {i}: material {m} rejected (rigid.js:144-214)")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def set_schedule(self, mode, order=None):
        o = np.ascontiguousarray(np.asarray(order if order is not None else [], dtype=np.int32))
        self.L.orc_set_schedule(self.h, int(mode), _ip(o), int(o.size))

    def set_body(self, i, pos=None, quat=None, vel=None, angvel=None):
        a = [None if x is None else _vec(x, n) for x, n in ((pos, 3), (quat, 4), (vel, 3), (angvel, 3))]
        self.L.orc_set_body(self.h, int(i), *[None if x is None else _dp(x) for x in a])

    def remove_bodies(self, doomed):
        """Solver.removeBodies (solver.js:261-300) for a boolean mask."""
        d = np.ascontiguousarray(np.asarray(doomed, dtype=np.int32))
        r = self.L.orc_remove_bodies(self.h, _ip(d), int(d.size))
        if r < 0:
            raise ValueError("mask length differs from the body count")
        return r

    def add_bodies(self, bodies):
        """Appends bodies (new Rigid(...) on a running solver); returns their ids."""
        ids = []
        for b in bodies:
            q = _vec(b.get("quat", (0, 0, 0, 1)), 4)
            w = _vec(b.get("angvel", (0, 0, 0)), 3)
            vel = _vec(b.get("vel", (0, 0, 0)), 3)
            pos = _vec(b["pos"], 3)
            if b["shape"] == "sphere":
                ids.append(self.L.orc_add_sphere(self.h, float(b["size"][0]), float(b["density"]),
                                                 float(b["friction"]), _dp(pos), _dp(vel), _dp(q), _dp(w)))
            else:
                size = _vec(b["size"], 3)
                ids.append(self.L.orc_add_box(self.h, _dp(size), float(b["density"]), float(b["friction"]),
                                              _dp(pos), _dp(vel), _dp(q), _dp(w)))
            self._set_material(ids[-1], b)
        return ids

    def record_pairs(self, on=True):
        self.L.orc_record_pairs(self.h, int(on))

    def step(self, n=1):
        """Run n steps; returns wall seconds."""
        return self.L.orc_step(self.h, int(n))

    @property
    def n(self):
        return self.L.orc_num_bodies(self.h)

    @property
    def n_dynamic(self):
        return self.L.orc_num_dynamic(self.h)

    def state(self):
        """(n, 16): p[3] q[4] v[3] w[3] vprev[3]."""
        out = np.zeros((self.n, 16), dtype=np.float64)
        self.L.orc_get_state(self.h, _dp(out))
        return out

    def format_state(self, step):
        """The frame of reference_main.cpp's trajectory dump (lines 126-140)
        for the current state, formatted by C's printf %.17g."""
        n = self.L.orc_format_state(self.h, int(step), None, 0)
        buf = C.create_string_buffer(n + 1)
        self.L.orc_format_state(self.h, int(step), buf, n + 1)
        return buf.value.decode()

    def mass(self):
        out = np.zeros((self.n, 8), dtype=np.float64)
        self.L.orc_get_mass(self.h, _dp(out))
        return out

    def forces(self):
        """(kind, a, b, nc) per live force in list order, and seq."""
        m = self.L.orc_num_forces(self.h)
        kab = np.zeros((m, 4), dtype=np.int32)
        seq = np.zeros(m, dtype=np.int64)
        self.L.orc_get_forces(self.h, _ip(kab), seq.ctypes.data_as(C.POINTER(C.c_longlong)))
        return kab, seq

    def manifold(self, idx):
        bm = np.zeros(10, dtype=np.float64)
        feat = np.zeros(8, dtype=np.uint32)
        con = np.zeros((8, 19), dtype=np.float64)
        nc = self.L.orc_get_manifold(self.h, int(idx), _dp(bm), feat.ctypes.data_as(C.POINTER(C.c_uint)),
                                     _dp(con))
        if nc < 0:
            raise IndexError(idx)
        ex = np.zeros(2, dtype=np.float64)
        bias = np.zeros(8, dtype=np.float64)
        self.L.orc_get_manifold_gpu(self.h, int(idx), _dp(ex), _dp(bias))
        return dict(basis=bm[:9].reshape(3, 3), mu=bm[9], features=feat[:nc], contacts=con[:nc],
                    mu_dynamic=ex[0], contact_distance=ex[1], bias=bias[:nc])

    def set_joint_anchor(self, joint, rA):
        """New rA of the joint created joint-th among joints and springs."""
        a = _vec(rA, 3)
        if self.L.orc_set_joint_anchor(self.h, int(joint), _dp(a)) < 0:
            raise IndexError(joint)

    def spring(self, idx):
        """Spring force idx: penalty, C, torn, peak strain, rest."""
        out = np.zeros(5, dtype=np.float64)
        if self.L.orc_get_spring(self.h, int(idx), _dp(out)) < 0:
            raise IndexError(idx)
        return out

    def joint(self, idx):
        out = np.zeros(25, dtype=np.float64)
        if self.L.orc_get_joint(self.h, int(idx), _dp(out)) < 0:
            raise IndexError(idx)
        return out

    def coloring(self):
        """Greedy colouring of the live force graph: (colorOf, offsets, entries)."""
        n = self.n
        col = np.zeros(n, dtype=np.int32)
        off = np.zeros(n + 2, dtype=np.int32)
        ent = np.zeros(max(n, 1), dtype=np.int32)
        k = self.L.orc_coloring(self.h, _ip(col), _ip(off), _ip(ent))
        return col, off[:k + 1], ent[:off[k]]

    def last_coloring(self):
        col = np.zeros(self.n, dtype=np.int32)
        k = self.L.orc_last_coloring(self.h, _ip(col))
        return col, k

    def pairs(self):
        m = self.L.orc_num_pairs(self.h)
        out = np.zeros((m, 2), dtype=np.int32)
        if m:
            self.L.orc_get_pairs(self.h, _ip(out))
        return out

    def collide(self, a, b):
        basis = np.zeros(9, dtype=np.float64)
        con = np.zeros((8, 7), dtype=np.float64)
        nc = self.L.orc_collide(self.h, int(a), int(b), _dp(basis), _dp(con))
        return nc, basis.reshape(3, 3), con[:nc]


def solve6(aLin, aAng, aCross, bLin, bAng):
    L = lib()
    a1, a2, a3 = (_vec(x, 9) for x in (aLin, aAng, aCross))
    b1, b2 = _vec(bLin, 3), _vec(bAng, 3)
    x1, x2 = np.zeros(3), np.zeros(3)
    L.orc_solve6(_dp(a1), _dp(a2), _dp(a3), _dp(b1), _dp(b2), _dp(x1), _dp(x2))
    return x1, x2


def render_poses(pos, quat, pos0, quat0, alpha):
    """Per body {p, 0, q} as float32 (n, 8): the pose a renderer draws at a
    fraction alpha of the last step, restating renderer.js:1327-1383 (alpha
    clamped to [0, 1]; below 1 the position is lerped from the step-start
    pose and the orientation nlerped along the shorter arc, in f64, then
    stored as f32). The renormalisation divides by sqrt(x^2+y^2+z^2+w^2)
    summed left to right where the renderer calls Math.hypot, which may
    differ from it in the last f64 bit."""
    a = 1.0 if math.isnan(alpha) else min(1.0, max(0.0, float(alpha)))
    p = np.array(pos, dtype=np.float64)
    q = np.array(quat, dtype=np.float64)
    if a < 1.0:
        p0 = np.asarray(pos0, dtype=np.float64)
        q0 = np.asarray(quat0, dtype=np.float64)
        p = p0 + (p - p0) * a
        d = q0[:, 0] * q[:, 0] + q0[:, 1] * q[:, 1] + q0[:, 2] * q[:, 2] + q0[:, 3] * q[:, 3]
        s = np.where(d < 0, -1.0, 1.0)[:, None]
        q = q0 * s + (q - q0 * s) * a
        ln = np.sqrt(q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1] + q[:, 2] * q[:, 2] + q[:, 3] * q[:, 3])
        q = q * (1.0 / np.where(ln != 0.0, ln, 1.0))[:, None]
    out = np.zeros((len(p), 8), np.float32)
    out[:, 0:3] = p.astype(np.float32)
    out[:, 4:8] = q.astype(np.float32)
    return out
