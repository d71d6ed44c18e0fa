"""The full-size configuration fixtures are the oracle's own output.

tests/golden/configs/<name>.json (make_config_vectors.py) pins the GPU's exact
mode at full size; here the oracle re-runs the first steps of each
configuration and must reproduce the committed fingerprints (the whole
trajectory for the small ones). ramp_fd.json (make_ramp_fd.py): the loss and
two of the central differences are recomputed.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import avbd
import pyoracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "configs")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def parts(name):
    S = avbd.scenes
    if name == "stack10":
        scene, params, meta = S.stack10()
        return [scene], params
    if name == "pyramid":
        scene, params, meta = S.pyramid()
        return [scene], params
    if name == "rain":
        scene, params, meta = S.rain()
        return [scene], params
    if name == "batched":
        scene, params, meta = S.batched()
        return meta["parts"], params
    scene, params, meta = S.ramp(64)
    return [scene], params


@pytest.mark.parametrize("name,steps", [("stack10", 300), ("ramp", 120), ("batched", 4), ("rain", 3),
                                        ("pyramid", 1)])
def test_fixture_is_the_oracle_trajectory(name, steps):
    doc = json.load(open(os.path.join(GOLDEN, name + ".json")))
    ps, params = parts(name)
    params = dict(params, colorOrder=1)
    worlds = [pyoracle.Oracle(p, params, schedule=pyoracle.SCHED_COLOR, color_order=1) for p in ps]
    for w in worlds:
        w.record_pairs(True)
    for s, rec in enumerate(doc["perStep"][:steps], start=1):
        base, pairs = 0, []
        for w in worlds:
            w.step(1)
            pairs.append(w.pairs() + base)
            base += w.n
        st = np.concatenate([w.state() for w in worlds], axis=0)[:, :13] + 0.0
        pl = np.concatenate(pairs[::-1], axis=0).astype(np.int32)
        assert sha(st) == rec["state"], f"{name} step {s}"
        assert len(pl) == rec["pairs"] and sha(pl) == rec["pairSha"], f"{name} step {s}: pairs"
        assert max(w.last_coloring()[1] for w in worlds) == rec["colors"]


def test_ramp_finite_differences_recompute():
    doc = json.load(open(os.path.join(GOLDEN, "ramp_fd.json")))
    scene, params, meta = avbd.scenes.ramp(64)
    target = np.array(meta["targets"])

    def loss(sc):
        o = pyoracle.Oracle(sc, params, schedule=pyoracle.SCHED_COLOR)
        o.step(meta["steps"])
        return float(((o.state()[2:, :3] - target[2:]) ** 2).sum())

    assert loss(scene) == doc["loss"]
    h = 1e-5
    for i in (4, 20):
        sp = dict(scene, bodies=[dict(b) for b in scene["bodies"]])
        sm = dict(scene, bodies=[dict(b) for b in scene["bodies"]])
        sp["bodies"][i]["friction"] += h
        sm["bodies"][i]["friction"] -= h
        assert (loss(sp) - loss(sm)) / (2 * h) == doc["fd"][str(i)]["mu"][repr(h)]
