"""GPU parity: the HIP step against the oracle and the reference's own vectors.

Exact precision (AVBD_PREC_EXACT) must reproduce, at every step, bit for bit
(post-stabilisation, the body-frame inertia and mid-run removals and additions
included):
  * the body state hashes the reference engine wrote (tests/golden, colour
    schedule: the reference Solver with its bodyOrder hook fed the greedy
    colouring taken after force initialisation, solver.js:225-233 and
    coloring.js:43-118);
  * the oracle's broad-phase pair list (A desc, B desc; broadphase.js:236-243);
  * the oracle's force list (kinds, bodies, contact counts, feature keys);
  * the oracle's colouring.
"""
import glob
import hashlib
import json
import os

import numpy as np
import pytest

import avbd
import pyoracle

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
COLOR_SCENES = sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(GOLDEN, "*_color.json")) +
                      glob.glob(os.path.join(GOLDEN, "*_prio.json")))
MAX_STEPS = 120


def gpu_state13(g):
    st = g.read_state()
    s = np.concatenate([st["pos"], st["quat"], st["vel"], st["angvel"]], axis=1)
    return s + 0.0


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def oracle_forces(o):
    kab, _ = o.forces()
    feats = np.zeros((len(kab), 8), np.uint32)
    for i, (kind, a, b, nc) in enumerate(kab):
        if kind == 0:
            f = o.manifold(i)["features"]
            feats[i, :len(f)] = f
    return kab, feats


@pytest.mark.parametrize("name", COLOR_SCENES)
def test_exact_bit_identical_to_reference(name):
    doc = json.load(open(os.path.join(GOLDEN, name + ".json")))
    print(f"[parity] {name}: {len(doc['scene']['bodies'])} bodies", flush=True)
    order = 1 if doc["schedule"] == "prio" else 0
    gpu = avbd.HipBackend(precision=avbd.PREC_EXACT)
    gpu.pack(doc["scene"], dict(doc["params"], colorOrder=order))
    ref = pyoracle.Oracle(doc["scene"], doc["params"], schedule=pyoracle.SCHED_COLOR, color_order=order)
    ref.record_pairs(True)
    cur = doc["scene"]
    for s, rec in enumerate(doc["perStep"][:MAX_STEPS], start=1):
        gpu.step(1)
        ref.step(1)
        gp, rp = gpu.read_pairs(), ref.pairs()
        assert np.array_equal(gp, rp), f"step {s}: pair lists differ ({len(gp)} vs {len(rp)})"
        gk, gf = gpu.read_forces()
        rk, rf = oracle_forces(ref)
        assert np.array_equal(gk, rk), f"step {s}: force lists differ"
        assert np.array_equal(gf, rf), f"step {s}: feature keys differ"
        gc, gn = gpu.read_coloring()
        rc, rn = ref.last_coloring()
        assert gn == rn and np.array_equal(gc, rc), f"step {s}: colourings differ"
        assert gc.tolist() == rec["colorOf"]
        st = gpu_state13(gpu)
        if sha(st) != rec["state"]:
            o = ref.state()[:, :13] + 0.0
            diff = np.abs(st - o)
            i, k = np.unravel_index(np.argmax(diff), diff.shape)
            pytest.fail(f"step {s}: state differs from the reference; max |diff| {diff.max():.3e} "
                        f"at body {i} component {k}")
        cur = apply_edits(gpu, ref, doc, s, cur)


def apply_edits(gpu, ref, doc, s, cur):
    """The fixture's topology edits after step s (Solver.removeBodies and new
    Rigid, solver.js:261-300): the oracle removes and appends; the GPU keeps
    the survivors' device state through avbd_repack with each body's old
    index. Returns the scene description after the edits."""
    for e in doc.get("edits", []):
        if e["step"] != s:
            continue
        n = len(cur["bodies"])
        gone = set(e.get("remove", []))
        keep = [i for i in range(n) if i not in gone]
        remap = {o: k for k, o in enumerate(keep)}
        joints = [dict(j, a=remap[j["a"]] if j["a"] >= 0 else -1, b=remap[j["b"]]) for j in cur["joints"]
                  if j["b"] in remap and (j["a"] < 0 or j["a"] in remap)]
        added = e.get("add", [])
        cur = dict(bodies=[cur["bodies"][i] for i in keep] + added, joints=joints)
        if gone:
            mask = np.zeros(n, dtype=np.int32)
            mask[sorted(gone)] = 1
            ref.remove_bodies(mask)
        if added:
            ref.add_bodies(added)
        gpu.repack(cur, keep + [-1] * len(added))
    return cur


@pytest.mark.parametrize("order", [0, 1], ids=["index", "priority"])
def test_exact_multi_workgroup_scene(order):
    # enough bodies that every kernel runs many workgroups (the golden scenes
    # fit in one); pairs, force lists, colouring and state against the oracle
    scene, params, _ = avbd.scenes.pyramid(layers=12)
    params = dict(params, colorOrder=order)
    gpu = avbd.HipBackend(precision=avbd.PREC_EXACT)
    gpu.pack(scene, params)
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR, color_order=order)
    ref.record_pairs(True)
    for s in range(1, 7):
        gpu.step(1)
        ref.step(1)
        assert np.array_equal(gpu.read_pairs(), ref.pairs()), f"step {s}: pairs"
        gk, gf = gpu.read_forces()
        rk, rf = oracle_forces(ref)
        assert np.array_equal(gk, rk) and np.array_equal(gf, rf), f"step {s}: forces"
        gc, gn = gpu.read_coloring()
        rc, rn = ref.last_coloring()
        assert gn == rn and np.array_equal(gc, rc), f"step {s}: colouring"
        st = gpu.read_state()
        o = ref.state()
        assert np.array_equal(st["pos"], o[:, :3]) and np.array_equal(st["quat"], o[:, 3:7]), f"step {s}: state"


def test_broad_phase_crowded_bodies_take_the_serial_path():
    # 150 overlapping spheres in a 1 m box: the later ones have more than the
    # 64 pairs a body's ranked list holds, so their lists come from the serial
    # path; the pair lists and the step still equal the oracle's
    r = avbd.scenes.lcg(17)
    bodies = [avbd.scenes.ground()]
    for _ in range(150):
        bodies.append(avbd.scenes.sphere(0.6, 1.0, 0.5, [r(), r(), 1.0 + r()]))
    scene = dict(bodies=bodies, joints=[])
    gpu = avbd.HipBackend(precision=avbd.PREC_EXACT)
    gpu.pack(scene)
    ref = pyoracle.Oracle(scene, {}, schedule=pyoracle.SCHED_COLOR)
    ref.record_pairs(True)
    for s in range(1, 4):
        gpu.step(1)
        ref.step(1)
        gp, rp = gpu.read_pairs(), ref.pairs()
        assert np.array_equal(gp, rp), f"step {s}: pairs"
        if s == 1:
            assert np.bincount(gp[:, 0]).max() > 64  # the serial path was taken
        st = gpu.read_state()
        assert np.array_equal(st["pos"], ref.state()[:, :3]), f"step {s}: state"


def test_exact_counters_and_manifold_records():
    doc = json.load(open(os.path.join(GOLDEN, "stack10_color.json")))
    gpu = avbd.HipBackend(precision=avbd.PREC_EXACT)
    gpu.pack(doc["scene"], doc["params"])
    ref = pyoracle.Oracle(doc["scene"], doc["params"], schedule=pyoracle.SCHED_COLOR)
    gpu.step(20)
    ref.step(20)
    c = gpu.counters()
    kab, _ = ref.forces()
    assert c["steps"] == 20 and c["bodies"] == 11 and c["dynamic_bodies"] == 10
    assert c["manifolds"] == int((kab[:, 0] == 0).sum())
    assert c["contacts"] == int(kab[:, 3].sum())
    assert c["overflow"] == 0
    for i in range(len(kab)):
        g = gpu.read_manifold(i)
        r = ref.manifold(i)
        assert np.array_equal(g["basis"], r["basis"]) and g["mu"] == r["mu"]
        # rA rB C0 k lambda, stick
        assert np.array_equal(g["contacts"][:, :15], r["contacts"][:, :15])
        assert np.array_equal(g["contacts"][:, 15], r["contacts"][:, 18])


@pytest.mark.parametrize("scene_name", ["pyramid12", "chain", "batched6"])
def test_exact_team_sweep_equals_one_lane_sweep(scene_name):
    # The exact mode's team sweep (a team of lanes per body, each contact's
    # contributions added in force order by one lane per matrix entry) and the
    # one-lane sweep (team 1, every body's system assembled by one lane) give
    # the same bits, step after step; scenes with joints mix the two (bodies
    # with joints take the one-lane path).
    if scene_name == "pyramid12":
        scene, params, _ = avbd.scenes.pyramid(layers=12)
    elif scene_name == "chain":
        doc = json.load(open(os.path.join(GOLDEN, "chain_color.json")))
        scene, params = doc["scene"], doc["params"]
    else:
        scene, params, _ = avbd.scenes.batched(num_scenes=6, seed=11)
    params = dict(params, colorOrder=1)
    team = avbd.HipBackend(precision=avbd.PREC_EXACT)
    lane = avbd.HipBackend(precision=avbd.PREC_EXACT, team=1)
    team.pack(scene, params)
    lane.pack(scene, params)
    for s in range(1, 16):
        team.step(1)
        lane.step(1)
        a, b = gpu_state13(team), gpu_state13(lane)
        assert np.array_equal(a, b), f"step {s}: max |diff| {np.abs(a - b).max():.3e}"
