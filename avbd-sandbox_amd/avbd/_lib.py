"""ctypes binding of the C ABI in include/avbd.h (libavbd_hip.so).

The library is the only compute path: if it is missing, or no GPU is present,
loading fails with an error; there is no CPU fallback.
"""
import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# AVBD_LIB may name another build of the same library (for instance one built
# with extra instrumentation); the default is the in-tree build.
LIB_PATH = os.environ.get("AVBD_LIB") or os.path.join(PKG_ROOT, "libavbd_hip.so")

AVBD_OK, AVBD_E_INVALID, AVBD_E_HIP, AVBD_E_CAPACITY, AVBD_E_STATE = 0, -1, -2, -3, -4
PREC_EXACT, PREC_F64, PREC_MIXED = 0, 1, 2
JOINT_BALL, JOINT_SPRING = 0, 1
SHAPE_BOX, SHAPE_SPHERE = 0, 1
STATIC_POSE_BODIES = 65536  # AVBD_STATIC_POSE_BODIES
NUM_PHASES = 6
PHASES = ("broad", "narrow", "topology", "colour", "solve", "total")

# Every symbol include/avbd.h declares.
SYMBOLS = (
    "avbd_abi_version", "avbd_default_params", "avbd_create", "avbd_destroy", "avbd_last_error",
    "avbd_set_params", "avbd_pack", "avbd_step", "avbd_sync", "avbd_read_state", "avbd_write_state",
    "avbd_read_pairs", "avbd_read_coloring", "avbd_read_forces", "avbd_read_manifold",
    "avbd_read_counters", "avbd_set_profiling", "avbd_read_profile", "avbd_read_kernel_stats",
    "avbd_autotune", "avbd_read_launch_shape", "avbd_set_launch_shape", "avbd_adjoint_forward", "avbd_adjoint_backward", "avbd_set_body_order",
    "avbd_upload_coloring", "avbd_export_poses", "avbd_export_poses_static", "avbd_pose_source", "avbd_repack",
    "avbd_read_joints",
    "avbd_read_step_start", "avbd_set_joint_anchor", "avbd_read_manifolds", "avbd_set_coloring_limits",
    "avbd_read_inertial_target", "avbd_read_contact_model",
)
K_PRIMAL, K_DUAL, K_NARROW, K_POSES = 0, 1, 2, 3

P = C.POINTER
d, i32, u32, vp = C.c_double, C.c_int32, C.c_uint32, C.c_void_p
pd, pi, pu = P(d), P(i32), P(u32)


class Config(C.Structure):
    _fields_ = [("device", i32), ("precision", i32), ("team", i32), ("manifold_cap", i32),
                ("pair_cap", i32), ("graphs", i32), ("adjoint_small", i32), ("adjoint_tape_mb", i32),
                ("scene_solve", i32)]


class Params(C.Structure):
    _fields_ = [("dt", d), ("gravity", d), ("beta_lin", d), ("beta_ang", d), ("alpha", d), ("gamma", d),
                ("penalty_min", d), ("penalty_max", d), ("collision_margin", d), ("stick_threshold", d),
                ("iterations", i32), ("post_stabilize", i32), ("rotated_inertia", i32),
                ("cached_contact_jacobians", i32), ("color_order", i32), ("contact_model", i32),
                ("contact_offset", d), ("rest_offset", d), ("restitution_threshold", d),
                ("contact_persistence", d)]


CONTACTS_CPU, CONTACTS_GPU = 0, 1

# Config.graphs: replay the solve loop of each step (0), whole steps (1), none (-1)
GRAPHS_SOLVE, GRAPHS_STEP, GRAPHS_NONE = 0, 1, -1


ORDER_INDEX, ORDER_PRIORITY, ORDER_CUSTOM = 0, 1, 2


class Scene(C.Structure):
    _fields_ = [("num_bodies", i32), ("shape", pi), ("size", pd), ("density", pd), ("friction", pd),
                ("pos", pd), ("quat", pd), ("vel", pd), ("angvel", pd), ("scene_of", pi),
                ("num_joints", i32), ("joint_a", pi), ("joint_b", pi), ("joint_ra", pd), ("joint_rb", pd),
                ("joint_stiff_lin", pd), ("joint_stiff_ang", pd), ("joint_fracture", pd), ("joint_kind", pi),
                ("static_friction", pd), ("dynamic_friction", pd), ("restitution", pd), ("contact_offset", pd),
                ("rest_offset", pd)]


class Counters(C.Structure):
    _fields_ = [("steps", C.c_int64), ("bodies", i32), ("dynamic_bodies", i32), ("pairs", i32),
                ("new_pairs", i32), ("manifolds", i32), ("contacts", i32), ("joints", i32), ("colors", i32),
                ("color_rounds", i32), ("overflow", i32)]


_lib = None


def lib(path=None):
    """Load libavbd_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"HIP library not built: {p} (run __graft_entry__.build())")
    L = C.CDLL(p)
    ctx = vp
    sig = {
        "avbd_abi_version": (C.c_int, []),
        "avbd_default_params": (None, [P(Params)]),
        "avbd_create": (C.c_int, [P(ctx), P(Config)]),
        "avbd_destroy": (C.c_int, [ctx]),
        "avbd_last_error": (C.c_char_p, [ctx]),
        "avbd_set_params": (C.c_int, [ctx, P(Params)]),
        "avbd_pack": (C.c_int, [ctx, P(Scene)]),
        "avbd_step": (C.c_int, [ctx, i32]),
        "avbd_sync": (C.c_int, [ctx]),
        "avbd_read_state": (C.c_int, [ctx, pd, pd, pd, pd, pd, i32]),
        "avbd_write_state": (C.c_int, [ctx, pd, pd, pd, pd, pd, i32]),
        "avbd_read_pairs": (C.c_int, [ctx, pi, i32, pi]),
        "avbd_read_coloring": (C.c_int, [ctx, pi, i32, pi]),
        "avbd_read_forces": (C.c_int, [ctx, pi, pu, i32, pi]),
        "avbd_read_manifold": (C.c_int, [ctx, i32, pd, pd, pi]),
        "avbd_read_counters": (C.c_int, [ctx, P(Counters)]),
        "avbd_set_profiling": (C.c_int, [ctx, i32]),
        "avbd_read_profile": (C.c_int, [ctx, P(C.c_float), i32, pi]),
        "avbd_read_kernel_stats": (C.c_int, [ctx, i32, pd, P(C.c_int64), pd]),
        "avbd_autotune": (C.c_int, [ctx, pi, P(C.c_float)]),
        "avbd_read_launch_shape": (C.c_int, [ctx, pi, pi, P(C.c_float)]),
        "avbd_set_launch_shape": (C.c_int, [ctx, i32, i32]),
        "avbd_adjoint_forward": (C.c_int, [ctx, i32, i32]),
        "avbd_adjoint_backward": (C.c_int, [ctx, pd, pd, pd]),
        "avbd_set_body_order": (C.c_int, [ctx, pi, i32]),
        "avbd_upload_coloring": (C.c_int, [ctx, pi, i32]),
        "avbd_export_poses": (C.c_int, [ctx, d, vp, i32, i32]),
        "avbd_export_poses_static": (C.c_int, [ctx, d, vp, i32]),
        "avbd_repack": (C.c_int, [ctx, P(Scene), pi]),
        "avbd_read_joints": (C.c_int, [ctx, pd, i32]),
        "avbd_read_step_start": (C.c_int, [ctx, pd, pd, i32]),
        "avbd_set_joint_anchor": (C.c_int, [ctx, i32, pd]),
        "avbd_pose_source": (C.c_int, [ctx, P(vp), P(vp), pi]),
        "avbd_read_manifolds": (C.c_int, [ctx, pi, pd, pd, i32, i32, pi, pi]),
        "avbd_set_coloring_limits": (C.c_int, [ctx, i32, i32]),
        "avbd_read_inertial_target": (C.c_int, [ctx, pd, pd, i32]),
        "avbd_read_contact_model": (C.c_int, [ctx, pd, pd, i32, i32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = L
    return L
