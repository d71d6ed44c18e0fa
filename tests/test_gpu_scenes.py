"""GPU behaviour beyond bit-exact parity: the f32/f64 variants against the
oracle, batched scenes, closed forms, and error paths.

Tolerances. Contact scenes switch discretely when a rounding-level difference
flips a tie in the reference's separating-axis choice (two face axes with the
same separation, as for a cube resting flat on the ground): the contact's
feature keys change and its warm start is dropped (manifold.js:241-286). From
that step the trajectories are different but equally valid, and resting stacks
at 10 iterations are chaotic (the reference's own stack collapses). The
variants are therefore held to a pose tolerance at every step up to the first
step whose force list or feature keys differ from the oracle's, and to
invariants (finite state, unit quaternions) after it:
  mixed (f32 contacts, f64 state and LDL^T):  5e-4 m / 5e-4 (quaternion)
  f64 with team reductions:                   1e-6
"""
import glob
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import avbd
import pyoracle
from avbd import AvbdError

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
SCENES = sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(GOLDEN, "*_color.json")))
TOL = {avbd.PREC_MIXED: 5e-4, avbd.PREC_F64: 1e-6}


def oracle_forces(o):
    kab, _ = o.forces()
    feats = np.zeros((len(kab), 8), np.uint32)
    for i, (kind, a, b, nc) in enumerate(kab):
        if kind == 0:
            f = o.manifold(i)["features"]
            feats[i, :len(f)] = f
    return kab, feats


def track(scene, params, steps, precision, order=0, team=16):
    """Steps GPU and oracle side by side: the first step whose force list or
    feature keys differ, the worst pose error before it (held to TOL), and
    the worst pose error after it."""
    gpu = avbd.HipBackend(precision=precision, team=team)
    gpu.pack(scene, dict(params, colorOrder=order))
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR, color_order=order)
    diverged, worst, after = None, 0.0, 0.0
    for s in range(1, steps + 1):
        gpu.step(1)
        ref.step(1)
        st = gpu.read_state()
        assert np.isfinite(st["pos"]).all() and np.isfinite(st["quat"]).all(), f"step {s}: non-finite state"
        qn = np.linalg.norm(st["quat"], axis=1)
        assert np.abs(qn - 1).max() < 1e-6, f"step {s}: quaternion norm drifted"
        o = ref.state()
        err = max(np.abs(st["pos"] - o[:, :3]).max(), np.abs(st["quat"] - o[:, 3:7]).max())
        if diverged is None:
            gk, gf = gpu.read_forces()
            rk, rf = oracle_forces(ref)
            if not (np.array_equal(gk, rk) and np.array_equal(gf, rf)):
                diverged = s
            else:
                worst = max(worst, err)
                assert err <= TOL[precision], f"step {s}: pose error {err:.3e} > {TOL[precision]:.0e}"
                continue
        after = max(after, err)
    return diverged, worst, after


# The first step whose force list or feature keys differ from the oracle's,
# per scene and precision, over the first 90 steps (None: none), as measured
# on MI355X with 16 lanes per body: a mode that diverges earlier fails. Every
# divergence comes from a tie in the separating-axis choice that rounding
# flips (collide.js:291-324); the f64-team precision meets none in 90 steps.
FIRST_DIVERGENCE = {
    "pyramid5_color": 2, "edits_color": 2, "stack10_color": 5, "poststab_color": 8, "springs_color": 83,
}
# After a divergence the trajectories are different valid ones; the pose
# difference to the oracle over the rest of the 90 steps is held to about ten
# times what was measured (pyramid5 1.9e-3, edits 9.2e-5, stack10 0.33 as the
# stack falls differently, poststab 3.5e-3, springs 1.3e-2; m or a quaternion
# component), which a blow-up or a body falling through the ground exceeds.
AFTER_BOUND = {"pyramid5_color": 0.02, "edits_color": 1e-3, "stack10_color": 2.0, "poststab_color": 0.05,
               "springs_color": 0.2}


def plank_on_cubes(n=4, pitch=1.25):
    """A plank resting on an n x n grid of unit cubes 0.25 m apart, on the
    ground. The plank touches every cube: n*n manifolds, about 4 n*n contact
    items against a cube's 8, so the team sweeps list it as a heavy body
    (more items than twice the team and twice the mean) and sweep it with a
    team twice as wide."""
    g = avbd.scenes
    bodies = [g.ground(0.6)]
    off = (n - 1) * pitch / 2
    for i in range(n):
        for j in range(n):
            bodies.append(g.box([1, 1, 1], 1.0, 0.6, [i * pitch - off, j * pitch - off, 0.5]))
    side = (n - 1) * pitch + 1.0
    bodies.append(g.box([side, side, 0.5], 2.0, 0.6, [0, 0, 1.25]))
    return dict(bodies=bodies, joints=[]), dict(gravity=-9.81, iterations=10)


@pytest.mark.parametrize("precision", [avbd.PREC_MIXED, avbd.PREC_F64], ids=["mixed", "f64"])
@pytest.mark.parametrize("team", [4, 8, 16])
def test_heavy_body_sweep_tracks_oracle(team, precision):
    scene, params = plank_on_cubes()
    diverged, worst, after = track(scene, params, 40, precision, team=team)
    print(f"plank team {team} precision {precision}: first divergence {diverged}, worst {worst:.2e}, after {after:.2e}")
    # measured on MI355X at teams 4, 8 and 16: f64 with team reductions keeps
    # the oracle's force lists for all 40 steps (pose error 4e-16); mixed
    # first flips a resting cube's face-axis tie at step 2 (error 9e-12 before
    # it) and stays within 2e-3 m of the oracle after it
    if precision == avbd.PREC_F64:
        assert diverged is None and worst < 1e-12
    else:
        assert diverged is None or diverged >= 2
        assert after < 0.02


def test_team_change_under_step_graphs_keeps_the_heavy_body_sweep():
    # f64 teams, whole steps replayed as graphs: 16 lanes per body (the plank
    # swept by 32), then 4 (the plank by 8) for the rest. The change of team
    # moves the heavy threshold; the steps after it read the new heavy-body
    # counts before their graphs are captured. The poses stay on the oracle's.
    scene, params = plank_on_cubes()
    gpu = avbd.HipBackend(precision=avbd.PREC_F64, team=16, graphs=avbd.GRAPHS_STEP)
    gpu.pack(scene, params)
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR, color_order=0)
    worst = 0.0
    for chunk, team in ((6, None), (1, 4), (13, None)):
        if team:
            gpu.set_launch_shape(team, 64)
        gpu.step(chunk)
        ref.step(chunk)
        st = gpu.read_state()
        o = ref.state()
        worst = max(worst, np.abs(st["pos"] - o[:, :3]).max(), np.abs(st["quat"] - o[:, 3:7]).max())
        gk, gf = gpu.read_forces()
        rk, rf = oracle_forces(ref)
        assert np.array_equal(gk, rk) and np.array_equal(gf, rf)
    assert gpu.launch_shape()[0] == 4
    assert worst < 1e-12, worst


@pytest.mark.parametrize("precision", [avbd.PREC_MIXED, avbd.PREC_F64], ids=["mixed", "f64"])
@pytest.mark.parametrize("name", SCENES)
def test_variant_tracks_oracle(name, precision):
    doc = json.load(open(os.path.join(GOLDEN, name + ".json")))
    steps = min(len(doc["perStep"]), 90)
    diverged, worst, after = track(doc["scene"], doc["params"], steps, precision)
    print(f"[{name}] precision {precision}: first divergence {diverged}, worst pose error before it {worst:.2e}, "
          f"after it {after:.2e}")
    floor = FIRST_DIVERGENCE.get(name) if precision == avbd.PREC_MIXED else None
    if floor is None:
        assert diverged is None, f"{name}: force lists diverged at step {diverged}"
    else:
        assert diverged is None or diverged >= floor, f"{name}: force lists diverged at step {diverged} < {floor}"
        assert after <= AFTER_BOUND[name], f"{name}: pose difference {after:.3e} after the divergence"


@pytest.mark.parametrize("order", [0, 1], ids=["index", "priority"])
def test_batched_scenes_equal_scenes_alone(order):
    scene, params, meta = avbd.scenes.batched(num_scenes=6, seed=11)
    steps = 30
    gpu = avbd.HipBackend(precision=avbd.PREC_EXACT)
    gpu.pack(scene, dict(params, colorOrder=order))
    gpu.step(steps)
    st = gpu.read_state()
    got = np.concatenate([st["pos"], st["quat"]], axis=1)
    base = 0
    for part in meta["parts"]:
        n = len(part["bodies"])
        ref = pyoracle.Oracle(part, params, schedule=pyoracle.SCHED_COLOR, color_order=order)
        ref.step(steps)
        assert np.array_equal(got[base:base + n], ref.state()[:, :7]), f"scene at body {base} differs"
        base += n
    c = gpu.counters()
    assert c["joints"] == len(scene["joints"]) and c["overflow"] == 0


@pytest.mark.parametrize("precision", [avbd.PREC_EXACT, avbd.PREC_F64, avbd.PREC_MIXED])
def test_free_fall_closed_form(precision):
    # analytic.mjs:43-56: BDF1 free fall is exact, z_n = z0 + g dt^2 n(n+1)/2
    scene = dict(bodies=[avbd.scenes.box([1, 1, 1], 1, 0.5, [0, 0, 100]),
                         avbd.scenes.sphere(1, 2, 0.5, [5, 0, 50], vel=[1.5, -0.5, 0])], joints=[])
    gpu = avbd.HipBackend(precision=precision)
    gpu.pack(scene, dict(gravity=-10.0))
    n = 120
    gpu.step(n)
    st = gpu.read_state()
    g, dt = -10.0, 1.0 / 60.0
    z = np.array([100.0, 50.0]) + g * dt * dt * (n * (n + 1) / 2)
    assert np.abs(st["pos"][:, 2] - z).max() / 100 < 1e-12
    assert abs(st["vel"][0, 2] - n * g * dt) < 1e-9
    # horizontal momentum survives untouched (analytic.mjs:62-72)
    assert abs(st["vel"][1, 0] - 1.5) < 1e-12 and abs(st["pos"][1, 0] - (5 + 1.5 * n * dt)) < 1e-9


def test_static_only_scene_and_counters():
    scene = dict(bodies=[avbd.scenes.ground(), avbd.scenes.box([1, 1, 1], 0, 0.5, [0, 0, 0.5])], joints=[])
    gpu = avbd.HipBackend(precision=avbd.PREC_MIXED)
    gpu.pack(scene)
    gpu.step(3)
    c = gpu.counters()
    assert c["pairs"] == 0 and c["manifolds"] == 0 and c["colors"] == 0 and c["dynamic_bodies"] == 0
    assert np.array_equal(gpu.read_state()["pos"][1], [0, 0, 0.5])


def test_error_paths():
    gpu = avbd.HipBackend(precision=avbd.PREC_MIXED)
    with pytest.raises(AvbdError) as e:
        gpu.pack(dict(bodies=[], joints=[]))
    assert e.value.code == -1
    with pytest.raises(AvbdError) as e:
        gpu.step(1)
    assert e.value.code == -4  # no scene packed
    bad = dict(bodies=[avbd.scenes.box([1, 1, 1], 1, 0.5, [0, 0, 1])],
               joints=[dict(a=0, b=0, rA=[0, 0, 0], rB=[0, 0, 0], kLin="inf", kAng=0, fracture="inf")])
    with pytest.raises(AvbdError) as e:
        gpu.pack(bad)
    assert e.value.code == -1
    with pytest.raises(AvbdError):
        gpu.set_params(colorOrder=7)


def test_pair_capacity_overflow_is_reported():
    scene, params, _ = avbd.scenes.pyramid(layers=6)
    gpu = avbd.HipBackend(precision=avbd.PREC_MIXED, pair_cap=16)
    gpu.pack(scene, params)
    with pytest.raises(AvbdError) as e:
        gpu.step(1)
    assert e.value.code == -3 and "capacity" in str(e.value)


def test_launch_through_write_state_matches_oracle():
    # the pyramid's impact: the parked sphere is launched at the apex by
    # overwriting its pose and velocity (the sandbox's pick-and-throw)
    scene, params, meta = avbd.scenes.pyramid(layers=4)
    L = meta["launch"]
    gpu = avbd.HipBackend(precision=avbd.PREC_EXACT)
    gpu.pack(scene, params)
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR)
    gpu.step(5)
    ref.step(5)
    st = gpu.read_state()
    pos, vel = st["pos"].copy(), st["vel"].copy()
    pos[L["body"]] = L["pos"]
    vel[L["body"]] = L["vel"]
    prev = st["prev_vel"].copy()
    prev[L["body"]] = L["vel"]
    gpu.write_state(pos=pos, vel=vel, prev_vel=prev)
    ref.set_body(L["body"], pos=L["pos"], vel=L["vel"])
    gpu.step(20)
    ref.step(20)
    g = gpu.read_state()
    o = ref.state()
    assert np.array_equal(g["pos"], o[:, :3]) and np.array_equal(g["quat"], o[:, 3:7])
    assert g["pos"][L["body"], 2] < L["pos"][2] - 1.0  # the sphere came down onto the apex


def test_autotune_picks_the_fastest_team():
    scene, params, _ = avbd.scenes.pyramid(layers=10)
    a = avbd.HipBackend(precision=avbd.PREC_MIXED)
    a.pack(scene, params)
    a.step(3)
    team, ms = a.autotune()
    print(f"autotune: team {team}, ms per sweep {ms}")
    assert team in (1, 4, 8, 16, 32, 64) and all(v > 0 for v in ms.values())
    assert ms[team] == min(ms.values())
    a.step(5)
    st = a.read_state()
    assert np.isfinite(st["pos"]).all() and np.isfinite(st["quat"]).all()
    e = avbd.HipBackend(precision=avbd.PREC_EXACT)
    e.pack(scene, params)
    with pytest.raises(AvbdError) as err:
        e.autotune()
    assert err.value.code == -4  # the exact mode has one team shape


def test_workgroup_size_changes_no_bits():
    # a body's update is one team's work whichever workgroup holds the team:
    # the per-colour sweep at 64, 128 and 256 lanes per workgroup, and the
    # one-launch batch at 128, 256 and 512 lanes per scene, give the same bits
    scene, params, _ = avbd.scenes.pyramid(layers=8)
    runs = []
    for blk in (64, 128, 256):
        g = avbd.HipBackend(precision=avbd.PREC_MIXED, team=16)
        g.pack(scene, params)
        g.set_launch_shape(16, blk)
        g.step(8)
        assert g.launch_shape()[:2] == (16, blk)
        runs.append(g.read_state())
    for st in runs[1:]:
        assert np.array_equal(st["pos"], runs[0]["pos"]) and np.array_equal(st["quat"], runs[0]["quat"])
    # the exact team sweep at each workgroup: the oracle's bits
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR)
    ref.step(8)
    for blk in (64, 128, 256):
        g = avbd.HipBackend(precision=avbd.PREC_EXACT)
        g.pack(scene, params)
        g.set_launch_shape(32, blk)
        g.step(8)
        assert g.launch_shape()[:2] == (32, blk)
        st = g.read_state()
        assert np.array_equal(st["pos"], ref.state()[:, :3]) and np.array_equal(st["quat"], ref.state()[:, 3:7])
    bscene, bparams, _ = avbd.scenes.batched(num_scenes=12)
    runs = []
    for blk in (128, 256, 512):
        g = avbd.HipBackend(precision=avbd.PREC_MIXED)
        g.pack(bscene, bparams)
        g.set_launch_shape(4, blk)
        g.step(6)
        assert g.launch_shape()[:2] == (4, blk)
        runs.append(g.read_state())
    for st in runs[1:]:
        assert np.array_equal(st["pos"], runs[0]["pos"]) and np.array_equal(st["quat"], runs[0]["quat"])
    with pytest.raises(AvbdError) as err:
        g.set_launch_shape(4, 64)  # not a batch workgroup
    assert err.value.code == -1


def test_autotune_tries_workgroup_sizes():
    scene, params, _ = avbd.scenes.pyramid(layers=8)
    a = avbd.HipBackend(precision=avbd.PREC_MIXED)
    a.pack(scene, params)
    a.step(2)
    team, ms = a.autotune()
    t, blk, trials = a.launch_shape()
    print(f"autotune: team {t}, block {blk}, trials {trials}")
    assert t == team and blk in (64, 128, 256)
    assert set(trials) == {(tm, b) for tm in (1, 4, 8, 16, 32, 64) for b in (64, 128, 256)}
    assert trials[(t, blk)] == min(trials.values())
    assert all(ms[tm] == min(trials[(tm, b)] for b in (64, 128, 256)) for tm in ms)
    bscene, bparams, _ = avbd.scenes.batched(num_scenes=16)
    b = avbd.HipBackend(precision=avbd.PREC_MIXED)
    b.pack(bscene, bparams)
    b.step(2)
    b.autotune()
    t, blk, trials = b.launch_shape()
    print(f"batch autotune: team {t}, block {blk}, trials {trials}")
    assert set(trials) == {(tm, g) for tm in (4, 8, 16) for g in (128, 256, 512)}
    assert trials[(t, blk)] == min(trials.values())
    b.step(3)
    assert np.isfinite(b.read_state()["pos"]).all()


def test_autotune_trials_restore_the_state():
    # spheres resting on the ground apart from each other: one contact item
    # per body, so every team shape sums the same terms exactly and a run
    # equals any other bit for bit; the trials must leave no trace
    bodies = [avbd.scenes.ground()]
    for i in range(24):
        bodies.append(avbd.scenes.sphere(1.0, 1.0, 0.5, [3.0 * (i % 6), 3.0 * (i // 6), 0.5 + 0.02 * (i % 5)],
                                         vel=[0.3 * (i % 3), 0.0, 0.0]))
    scene = dict(bodies=bodies, joints=[])
    a = avbd.HipBackend(precision=avbd.PREC_MIXED)
    a.pack(scene, dict(iterations=10))
    a.step(3)
    team, _ = a.autotune()
    a.step(5)
    b = avbd.HipBackend(precision=avbd.PREC_MIXED, team=team)
    b.pack(scene, dict(iterations=10))
    b.step(9)
    sa, sb = a.read_state(), b.read_state()
    assert np.array_equal(sa["pos"], sb["pos"]) and np.array_equal(sa["quat"], sb["quat"])
    assert a.counters()["contacts"] == 24


@pytest.mark.parametrize("graphs", [0, -1], ids=["graph", "direct"])
def test_graph_replay_equals_direct_launches(graphs):
    scene, params, _ = avbd.scenes.pyramid(layers=8)
    params = dict(params, colorOrder=1)
    g = avbd.HipBackend(precision=avbd.PREC_EXACT, graphs=graphs)
    g.pack(scene, params)
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR, color_order=1)
    g.step(12)
    ref.step(12)
    st = g.read_state()
    o = ref.state()
    assert np.array_equal(st["pos"], o[:, :3]) and np.array_equal(st["quat"], o[:, 3:7])


def test_unchanged_force_graph_confirms_the_colouring_in_one_round():
    # a settled stack keeps its contact set from step to step: the colouring
    # is confirmed by one relaxation round and still equals the oracle's
    scene, params, _ = avbd.scenes.stack10()
    gpu = avbd.HipBackend(precision=avbd.PREC_EXACT)
    gpu.pack(scene, params)
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR)
    one_round = 0
    for s in range(1, 151):
        gpu.step(1)
        ref.step(1)
        gc, gn = gpu.read_coloring()
        rc, rn = ref.last_coloring()
        assert gn == rn and np.array_equal(gc, rc), f"step {s}: colourings differ"
        one_round += gpu.counters()["color_rounds"] == 1
    assert np.array_equal(gpu.read_state()["pos"], ref.state()[:, :3])
    print(f"[colour] {one_round} of 150 steps confirmed in one round", flush=True)
    assert one_round > 0


def test_capacity_regrows_at_the_next_repack():
    # backend.js:1683-1739: an overflow is reported and the next repack grows
    # the arena; the step that overflowed left the state untouched, so the
    # scene continues bit for bit with the oracle
    scene, params, _ = avbd.scenes.pyramid(layers=4, speed=0.0)
    gpu = avbd.HipBackend(precision=avbd.PREC_EXACT, pair_cap=24)
    gpu.pack(scene, params)
    with pytest.raises(AvbdError) as e:
        gpu.step(1)
    assert e.value.code == -3
    gpu.repack(scene, list(range(len(scene["bodies"]))))
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR)
    for s in range(1, 21):
        gpu.step(1)
        ref.step(1)
        assert np.array_equal(gpu.read_state()["pos"], ref.state()[:, :3]), f"step {s}"


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_js_step_collects_errors_and_regrows(tmp_path):
    scene, params, _ = avbd.scenes.pyramid(layers=4, speed=0.0)
    path = tmp_path / "scene.json"
    path.write_text(json.dumps({"scene": scene, "params": params}))
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    url = "file://" + os.path.join(root, "avbd-sandbox_amd", "js")
    js = ("import { readFileSync } from 'fs';"
          f"import {{ HipBackend }} from '{url}/hip_backend.mjs';"
          f"import {{ fromDescription }} from '{url}/scene_graph.mjs';"
          f"const d = JSON.parse(readFileSync({json.dumps(str(path))}, 'utf8'));"
          "const solver = fromDescription(d.scene, d.params);"
          "const b = new HipBackend({ precision: 'exact', pairCap: 24 });"
          "(async () => { await b.init();"
          " for (let s = 0; s < 21; s++) b.step(solver);"
          " await b.syncToCPU(solver);"
          " console.log(JSON.stringify({ errors: b.gpuErrors, pos: solver.bodies.map((x) => Array.from(x.positionLin)) }));"
          "})().catch((e) => { console.error(e); process.exit(1); });")
    script = tmp_path / "grow.mjs"
    script.write_text(js)
    r = subprocess.run([shutil.which("node"), str(script)], capture_output=True, text=True, cwd=root, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    assert len(out["errors"]) == 1 and "capacity" in out["errors"][0]
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR)
    ref.step(20)  # 21 calls, one of which overflowed
    assert np.array_equal(np.array(out["pos"]), ref.state()[:, :3])


@pytest.mark.parametrize("precision", [avbd.PREC_EXACT, avbd.PREC_F64, avbd.PREC_MIXED], ids=["exact", "f64", "mixed"])
def test_one_launch_batch_solve_equals_per_colour_launches(precision):
    # A batch's solve as one launch with a workgroup per scene (colours
    # separated by workgroup barriers) against one launch per colour and
    # iteration: the same per-body updates in the same colour order. Exact:
    # the same bits. The team modes sum each body's contacts over a team of
    # lanes whose width the two paths choose separately, so they agree to
    # rounding until a tie flips a contact (the batch's stacks flip at step 3
    # in either path), and stay finite and close after.
    scene, params, meta = avbd.scenes.batched(num_scenes=20, seed=5)
    params = dict(params, colorOrder=1)
    one = avbd.HipBackend(precision=precision, team=8)
    per = avbd.HipBackend(precision=precision, team=8, scene_solve=-1)
    one.pack(scene, params)
    per.pack(scene, params)
    for s in range(1, 41):
        one.step(1)
        per.step(1)
        a, b = one.read_state(), per.read_state()
        if precision == avbd.PREC_EXACT:
            for k in ("pos", "quat", "vel", "angvel"):
                assert np.array_equal(a[k], b[k]), f"step {s}: {k}"
        else:
            assert np.isfinite(a["pos"]).all()
            if s <= 2:
                assert np.abs(a["pos"] - b["pos"]).max() <= (1e-4 if precision == avbd.PREC_MIXED else 1e-9)
    assert one.counters()["overflow"] == 0


@pytest.mark.parametrize("which", ["pyramid12", "rain3000"])
def test_mixed_precision_repeats_bit_for_bit_with_a_fixed_team(which):
    # With the lanes per body fixed (not autotuned by timing), every reduction
    # of the mixed precision has one order, and two runs give the same bits
    # (the run-to-run variation of a benchmark line comes from the autotuned
    # team shape alone)
    if which == "pyramid12":
        scene, params, _ = avbd.scenes.pyramid(layers=12)
    else:
        scene, params, _ = avbd.scenes.rain(n=3000)
    params = dict(params, colorOrder=1)
    runs = []
    for _ in range(2):
        g = avbd.HipBackend(precision=avbd.PREC_MIXED, team=16)
        g.pack(scene, params)
        out = []
        for _ in range(40):
            g.step(1)
            st = g.read_state()
            out.append(np.concatenate([st["pos"], st["quat"], st["vel"], st["angvel"]], axis=1))
        runs.append(out)
    for s, (a, b) in enumerate(zip(*runs), start=1):
        assert np.array_equal(a, b), f"step {s}: runs differ by {np.abs(a - b).max():.3e}"
