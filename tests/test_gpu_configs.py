"""The five benchmark configurations at their full sizes, on the GPU.

Exact precision against the oracle. tests/golden/configs/<name>.json holds the
f64 oracle's trajectory of each configuration (make_config_vectors.py): per
step, the SHA-256 of every body's state, the broad-phase pair list's count and
SHA-256, and the colour count. The exact mode must reproduce every one of
them, bit for bit, at every step the fixture holds: the 10-cube stack for its
300 steps, the 51,039-cube pyramid through the sphere's launch at step 60 and
the impact, the 50,000-body rain, the 256 batched scenes for their 300 steps
and the 64-box ramp for its 120 steps.

The throughput precisions against the exact mode. With exact mode pinned to the
oracle, it stands in for the oracle over whole windows. The mixed (f32
contacts, f64 state and LDL^T) and f64-team modes run in lockstep beside it;
per configuration the test measures the first step whose force list or
contact feature keys differ from exact mode's (a tie in the separating-axis
choice that rounding flips, collide.js:291-324; from there the trajectories
are different but equally valid ones, and stacks at these iteration counts
are chaotic), the largest pose error before it, and the largest pose error
over the window. Each is held to a per-configuration bound (VARIANT below):
a mode that diverged at step 1, or drifted, fails.

Quaternions are compared as |q_a - q_b| after aligning signs; pose error is
the larger of the position error (m) and that quaternion error.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import avbd

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(__file__)
CONFIGS = os.path.join(HERE, "golden", "configs")
REPORT = os.environ.get("AVBD_CONFIG_REPORT")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load(name):
    path = os.path.join(CONFIGS, name + ".json")
    if not os.path.exists(path):
        pytest.skip(f"no oracle trajectory for {name}")
    with open(path) as f:
        return json.load(f)


def config_scene(name):
    """The configuration's scene as the bench packs it (batched: one scene
    with scene ids)."""
    S = avbd.scenes
    if name == "stack10":
        scene, params, meta = S.stack10()
    elif name == "pyramid":
        scene, params, meta = S.pyramid()
    elif name == "rain":
        scene, params, meta = S.rain()
    elif name == "batched":
        scene, params, meta = S.batched()
    else:
        scene, params, meta = S.ramp(64)
    return scene, dict(params, colorOrder=1), meta


def launch(gpu, spec):
    """Sets one body's position and velocity (and previous velocity), angular
    velocity 0: the pyramid's strike (scenes.pyramid meta['launch'])."""
    st = gpu.read_state()
    b = spec["body"]
    st["pos"][b] = spec["pos"]
    st["vel"][b] = spec["vel"]
    st["prev_vel"][b] = spec["vel"]
    st["angvel"][b] = 0.0
    gpu.write_state(st["pos"], st["quat"], st["vel"], st["angvel"], st["prev_vel"])


def state13(gpu):
    st = gpu.read_state()
    return np.concatenate([st["pos"], st["quat"], st["vel"], st["angvel"]], axis=1) + 0.0


@pytest.mark.parametrize("name", ["stack10", "ramp", "batched", "rain", "pyramid"])
def test_exact_reproduces_the_oracle_at_full_size(name):
    doc = load(name)
    scene, params, meta = config_scene(name)
    assert doc["bodies"] == len(scene["bodies"])
    print(f"[configs] exact {name}: {len(scene['bodies'])} bodies, {doc['steps']} steps", flush=True)
    gpu = avbd.HipBackend(precision=avbd.PREC_EXACT)
    gpu.pack(scene, params)
    spec = doc.get("launch")
    for s, rec in enumerate(doc["perStep"], start=1):
        gpu.step(1)
        pairs = gpu.read_pairs()
        assert len(pairs) == rec["pairs"] and sha(pairs) == rec["pairSha"], f"{name} step {s}: pair list"
        _, k = gpu.read_coloring()
        assert k == rec["colors"], f"{name} step {s}: {k} colours, oracle {rec['colors']}"
        assert sha(state13(gpu)) == rec["state"], f"{name} step {s}: body state differs from the oracle's"
        if spec is not None and s == spec["step"]:
            launch(gpu, spec)
    c = gpu.counters()
    assert c["overflow"] == 0
    if name == "pyramid" and doc["steps"] > spec["step"]:
        # the strike reached the pile within the fixture's window
        assert c["manifolds"] > 0


def pose_error(a, b, dyn):
    dp = np.abs(a["pos"][dyn] - b["pos"][dyn]).max(initial=0.0)
    qa, qb = a["quat"][dyn], b["quat"][dyn]
    sgn = np.where(np.sum(qa * qb, axis=1) < 0, -1.0, 1.0)[:, None]
    dq = np.abs(qa - sgn * qb).max(initial=0.0)
    return max(dp, dq)


def lambda_error(a, b):
    """Largest difference of any contact multiplier (normal and friction
    rows) between two contexts holding the same force list, and the largest
    multiplier (N: lambda is a force, manifold.js:447-491)."""
    _, _, ca = a.read_manifolds()
    _, _, cb = b.read_manifolds()
    if len(ca) == 0:
        return 0.0, 0.0
    return float(np.abs(ca[:, 12:15] - cb[:, 12:15]).max()), float(np.abs(ca[:, 12:15]).max())


def same_forces(a, b):
    ka, fa = a.read_forces()
    kb, fb = b.read_forces()
    return np.array_equal(ka, kb) and np.array_equal(fa, fb)


# Per configuration and precision: (steps compared, first force-list
# divergence no earlier than, pose error before it at most, pose error over
# the whole window at most, largest contact-multiplier difference before the
# divergence relative to the largest multiplier at most). The floors and bounds are the measured values
# (MI355X) with margin; the measured ones are tabulated in the design notes.
# floor None: no divergence within the window. bound None: after the first
# divergence the trajectories are different valid ones and only invariants are
# checked: the batched chains are chaotic pendulums (the check is that their
# sockets stay as closed as exact mode's), and in the rain the oracle's own
# trajectory ejects light bodies squeezed between heavy ones (densities
# 1:1000) at hundreds of m/s, so any difference there grows without bound.
VARIANT = {
    ("stack10", "mixed"): (300, 4, 1e-9, 5.0, 1e-6),
    ("stack10", "f64"): (300, None, 1e-10, 1e-10, 1e-10),
    ("ramp", "mixed"): (120, None, 5e-5, 5e-5, 1e-4),
    ("ramp", "f64"): (120, None, 1e-10, 1e-10, 1e-10),
    ("batched", "mixed"): (300, 3, 1e-8, None, 1e-6),
    ("batched", "f64"): (300, 3, 1e-12, None, 1e-12),
    ("pyramid", "mixed"): (120, 2, 1e-8, 2e-2, 1e-6),
    ("pyramid", "f64"): (120, 4, 1e-6, 2e-2, 5e-5),
    ("rain", "mixed"): (150, 10, 1e-5, None, 1e-5),
    ("rain", "f64"): (150, 30, 2e-2, None, 5e-4),
}


def qrot(q, v):
    """Rotates v (n, 3) by unit quaternions q (n, 4), xyzw."""
    u, w = q[:, :3], q[:, 3:4]
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def socket_gaps(st, joints):
    """Distance between the two anchors of every ball joint (m)."""
    if not joints:
        return np.zeros(0)
    a = np.array([j["a"] for j in joints])
    b = np.array([j["b"] for j in joints])
    rA = np.array([j["rA"] for j in joints], float)
    rB = np.array([j["rB"] for j in joints], float)
    pa = np.where((a >= 0)[:, None], st["pos"][a] + qrot(st["quat"][a], rA), rA)
    pb = st["pos"][b] + qrot(st["quat"][b], rB)
    return np.linalg.norm(pa - pb, axis=1)


@pytest.mark.parametrize("precision", ["mixed", "f64"])
@pytest.mark.parametrize("name", ["stack10", "ramp", "batched", "pyramid", "rain"])
def test_throughput_precisions_track_exact_at_full_size(name, precision):
    steps, floor, tol_before, bound, tol_lambda = VARIANT[(name, precision)]
    scene, params, meta = config_scene(name)
    spec = meta.get("launch")
    print(f"[configs] {precision} vs exact {name}: {len(scene['bodies'])} bodies, {steps} steps", flush=True)
    prec = {"mixed": avbd.PREC_MIXED, "f64": avbd.PREC_F64}[precision]
    ex = avbd.HipBackend(precision=avbd.PREC_EXACT)
    va = avbd.HipBackend(precision=prec)
    ex.pack(scene, params)
    va.pack(scene, params)
    dyn = np.array([b["density"] > 0 for b in scene["bodies"]])
    diverged, before, worst, worst_step = None, 0.0, 0.0, 0
    lam_err, lam_max = 0.0, 0.0
    per_step = []
    for s in range(1, steps + 1):
        ex.step(1)
        va.step(1)
        a, b = ex.read_state(), va.read_state()
        assert np.isfinite(b["pos"]).all() and np.isfinite(b["quat"]).all(), f"step {s}: non-finite state"
        assert np.abs(np.linalg.norm(b["quat"], axis=1) - 1).max() < 1e-6, f"step {s}: quaternion norm drifted"
        err = pose_error(a, b, dyn)
        per_step.append(err)
        if err > worst:
            worst, worst_step = err, s
        if diverged is None:
            if same_forces(ex, va):
                before = max(before, err)
                dl, lm = lambda_error(ex, va)
                lam_err, lam_max = max(lam_err, dl), max(lam_max, lm)
            else:
                diverged = s
        if spec is not None and s == spec["step"]:
            launch(ex, spec)
            launch(va, spec)
    gap_ex, gap_va = socket_gaps(a, scene["joints"]), socket_gaps(b, scene["joints"])
    rec = dict(config=name, precision=precision, steps=steps, first_divergence=diverged,
               error_before_divergence=before, worst_error=worst, worst_step=worst_step,
               lambda_error_before_divergence=lam_err, largest_lambda=lam_max,
               socket_gap=dict(exact=float(gap_ex.max(initial=0.0)), variant=float(gap_va.max(initial=0.0))),
               error_at=dict((str(k), per_step[k - 1]) for k in (10, 30, 60, 90, 120, 150, 300) if k <= steps))
    print("[configs] " + json.dumps(rec), flush=True)
    if REPORT:
        with open(REPORT, "a") as f:
            f.write(json.dumps(rec) + "\n")
    if floor is None:
        assert diverged is None, f"{name} {precision}: force lists diverged at step {diverged}"
    else:
        assert diverged is None or diverged >= floor, f"{name} {precision}: force lists diverged at step {diverged}"
    assert before <= tol_before, f"{name} {precision}: pose error {before:.3e} before the first divergence"
    assert lam_err <= tol_lambda * max(lam_max, 1.0), f"{name} {precision}: lambda error {lam_err:.3e} of {lam_max:.3e}"
    if bound is not None:
        assert worst <= bound, f"{name} {precision}: pose error {worst:.3e} at step {worst_step}"
    if len(gap_ex):
        assert gap_va.max() <= 2.0 * gap_ex.max() + 1e-3, f"{name} {precision}: socket gap {gap_va.max():.3e}"
