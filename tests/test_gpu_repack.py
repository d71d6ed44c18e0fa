"""Topology edits that keep state (avbd_repack; GpuBackend.pack's identity-
preserving repack, backend.js:662-1301) against the oracle's
Solver.removeBodies (solver.js:261-300) followed by appended bodies.

Exact precision: after an erase and append the GPU continues bit for bit
with the oracle (states, force lists with feature keys, colourings), with
contacts, joints (one of them fractured) and spheres. Mixed precision: the
reference's own contract for an index-shifting erase of a settled pile
(gpu_headless.mjs:1478-1535): the flush moves nothing, every manifold's
penalties and multipliers come through unchanged, and the pile does not move
by 0.02 m in the 24 steps after."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import avbd
import pyoracle
from avbd import AvbdError, sceneio
from avbd import scenes as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def state13(g):
    st = g.read_state()
    return np.concatenate([st["pos"], st["quat"], st["vel"], st["angvel"]], axis=1) + 0.0


def oracle_forces(o):
    kab, _ = o.forces()
    feats = np.zeros((len(kab), 8), np.uint32)
    for i, (kind, a, b, nc) in enumerate(kab):
        if kind == 0:
            f = o.manifold(i)["features"]
            feats[i, :len(f)] = f
    return kab, feats


def edit(scene, doomed, new_bodies):
    """The scene after erasing `doomed` and appending `new_bodies`, with each
    body's old index (-1 for new) and the surviving joints renumbered."""
    n = len(scene["bodies"])
    remap = np.full(n, -1)
    keep = [i for i in range(n) if i not in doomed]
    remap[keep] = np.arange(len(keep))
    bodies = [scene["bodies"][i] for i in keep] + list(new_bodies)
    joints = []
    for j in scene.get("joints", []):
        a, b = j["a"], j["b"]
        if remap[b] < 0 or (a >= 0 and remap[a] < 0):
            continue
        joints.append(dict(j, a=int(remap[a]) if a >= 0 else -1, b=int(remap[b])))
    old_index = keep + [-1] * len(new_bodies)
    return dict(bodies=bodies, joints=joints), old_index


def check_lockstep(gpu, ref, steps, label):
    for s in range(1, steps + 1):
        gpu.step(1)
        ref.step(1)
        gk, gf = gpu.read_forces()
        rk, rf = oracle_forces(ref)
        assert np.array_equal(gk, rk), f"{label} step {s}: force lists differ"
        assert np.array_equal(gf, rf), f"{label} step {s}: feature keys differ"
        gc, gn = gpu.read_coloring()
        rc, rn = ref.last_coloring()
        assert gn == rn and np.array_equal(gc, rc), f"{label} step {s}: colourings differ"
        g, o = state13(gpu), ref.state()[:, :13] + 0.0
        assert np.array_equal(g, o), f"{label} step {s}: max |diff| {np.abs(g - o).max():.3e}"


def test_erase_and_append_contacts_exact():
    scene, params, _ = S.pyramid(layers=6, speed=0.0)
    print(f"[repack] pyramid 6 layers: {len(scene['bodies'])} bodies", flush=True)
    gpu = avbd.HipBackend(precision=avbd.PREC_EXACT)
    gpu.pack(scene, params)
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR)
    check_lockstep(gpu, ref, 15, "before")
    doomed = [3, 20, 21, 40]
    new = [S.box([1, 1, 1], 1.0, 0.6, [0.2, 0.1, 8.0]), S.sphere(0.8, 2.0, 0.5, [-1.0, 0.5, 9.5], vel=[0, 0, -3]),
           S.box([0.5, 0.7, 0.9], 1.0, 0.4, [1.5, -0.5, 7.0], angvel=[0.5, 0.2, 0.0])]
    scene2, old_index = edit(scene, doomed, new)
    gpu.repack(scene2, old_index)
    mask = np.zeros(ref.n, bool)
    mask[doomed] = True
    ref.remove_bodies(mask)
    ref.add_bodies(new)
    assert gpu.n == ref.n == len(scene2["bodies"])
    assert np.array_equal(state13(gpu), ref.state()[:, :13] + 0.0)
    check_lockstep(gpu, ref, 40, "after")


def test_erase_with_joints_and_a_fracture_exact():
    # the reference parity harness's rope with hard angular rows; joint 2
    # gets a fracture threshold on ||lambda_ang|| (joint.js:307-316) that the
    # sagging rope exceeds at step 19, before the edit
    scene = sceneio.read_scene(os.path.join(ROOT, "tests", "golden", "scenes", "rope_angular.scene"))
    scene["joints"][2] = dict(scene["joints"][2], fracture=30.0)
    params = {}
    gpu = avbd.HipBackend(precision=avbd.PREC_EXACT)
    gpu.pack(scene, params)
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR)
    check_lockstep(gpu, ref, 30, "before")
    # a broken joint leaves the force list at the next step (joint.js:159)
    live_joints = int((ref.forces()[0][:, 0] == 1).sum())
    print(f"[repack] rope: {live_joints} of {len(scene['joints'])} joints live before the edit", flush=True)
    assert live_joints == len(scene["joints"]) - 1
    doomed = [8]  # a middle link: the joints on either side of it go
    scene2, old_index = edit(scene, doomed, [S.box([1, 1, 1], 1.0, 0.5, [3.0, 0.0, -15.0])])
    gpu.repack(scene2, old_index)
    mask = np.zeros(ref.n, bool)
    mask[doomed] = True
    ref.remove_bodies(mask)
    ref.add_bodies(scene2["bodies"][-1:])
    check_lockstep(gpu, ref, 30, "after")


def test_rejected_edits_leave_the_scene_running():
    bodies, joints = S.chain(6)
    scene, params = dict(bodies=bodies, joints=joints), dict(gravity=-9.81)
    gpu = avbd.HipBackend(precision=avbd.PREC_EXACT)
    gpu.pack(scene, params)
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR)
    check_lockstep(gpu, ref, 5, "before")
    n = len(bodies)
    swapped = list(range(n))
    swapped[2], swapped[3] = 3, 2
    with pytest.raises(AvbdError) as e:
        gpu.repack(scene, swapped)  # survivors out of order
    assert e.value.code == avbd._lib.AVBD_E_INVALID
    extra = dict(scene, joints=joints + [dict(joints[0], a=3, b=5)])
    with pytest.raises(AvbdError) as e:
        gpu.repack(extra, list(range(n)))  # a joint added by a repack
    assert e.value.code == avbd._lib.AVBD_E_INVALID
    with pytest.raises(AvbdError):
        gpu.repack(scene, [0] * (n - 1) + [n])  # out of range
    check_lockstep(gpu, ref, 10, "after rejections")


def test_index_shifting_erase_keeps_a_settled_pile_mixed():
    scene, params, _ = S.pyramid(layers=8, speed=0.0)
    # a physically irrelevant body at index 1, whose erase shifts every
    # other body's index by one (the reference's removablePrefix)
    scene["bodies"].insert(1, S.box([0.5, 0.5, 0.5], 1.0, 0.5, [60.0, 60.0, 0.25]))
    gpu = avbd.HipBackend(precision=avbd.PREC_MIXED)
    gpu.pack(scene, params)
    gpu.step(90)
    before = gpu.read_state()
    kab, _ = gpu.read_forces()
    mans = {(int(a), int(b)): gpu.read_manifold(i) for i, (k, a, b, nc) in enumerate(kab) if k == 0}
    scene2, old_index = edit(scene, [1], [])
    gpu.repack(scene2, old_index)
    after = gpu.read_state()
    for k in ("pos", "quat", "vel", "angvel", "prev_vel"):
        assert np.array_equal(after[k], np.delete(before[k], 1, axis=0)), k
    kab2, _ = gpu.read_forces()
    assert len(kab2) == len([1 for (a, b) in mans if a != 1 and b != 1])
    for i, (k, a, b, nc) in enumerate(kab2):
        old = mans[(int(a) + (a >= 1), int(b) + (b >= 1))]
        new = gpu.read_manifold(i)
        assert np.array_equal(old["contacts"], new["contacts"]) and np.array_equal(old["basis"], new["basis"])
    worst = 0.0
    for _ in range(24):
        gpu.step(1)
        worst = max(worst, np.abs(gpu.read_state()["pos"] - after["pos"]).max())
    print(f"[repack] settled pile: worst move {worst:.2e} m over 24 steps", flush=True)
    assert worst < 0.02


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_js_host_keeps_state_across_an_erase():
    # HipBackend.pack(solver) after Solver.removeBodies and a new Rigid: the
    # identity-preserving path, bit for bit with the oracle
    path = os.path.join(ROOT, "tests", "golden", "scenes", "rope_angular.scene")
    r = subprocess.run([shutil.which("node"), "tests/js/run_scene.mjs", "repack", path, "20", "exact"],
                       capture_output=True, text=True, cwd=ROOT, timeout=120)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    assert d["kept"] is True
    ref = pyoracle.Oracle(path, schedule=pyoracle.SCHED_COLOR)
    ref.step(20)
    mask = np.zeros(ref.n, bool)
    mask[8] = True
    ref.remove_bodies(mask)
    ref.add_bodies([S.box([1, 1, 1], 1, 0.5, [3, 0, -15])])
    ref.step(20)
    assert d["frame"] == ref.format_state(40)


def test_erase_with_springs_exact():
    # the spring fixture (reference engine goldens): a link of the stiffness-
    # ratio chain is erased after the tearable spring tore (step 12), so the
    # springs on either side of it go and the torn one stays, inactive
    doc = json.load(open(os.path.join(ROOT, "tests", "golden", "springs_color.json")))
    scene, params = doc["scene"], doc["params"]
    gpu = avbd.HipBackend(precision=avbd.PREC_EXACT)
    gpu.pack(scene, params)
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR)
    check_lockstep(gpu, ref, 20, "before")
    doomed = [6]
    scene2, old_index = edit(scene, doomed, [S.box([1, 1, 1], 1.0, 0.5, [0.0, 4.0, 6.0])])
    gpu.repack(scene2, old_index)
    mask = np.zeros(ref.n, bool)
    mask[doomed] = True
    ref.remove_bodies(mask)
    ref.add_bodies(scene2["bodies"][-1:])
    check_lockstep(gpu, ref, 30, "after")
