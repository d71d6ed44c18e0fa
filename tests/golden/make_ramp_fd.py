"""Writes tests/golden/configs/ramp_fd.json: f64 central finite differences of
the oracle for the differentiable configuration (64 boxes on a 30-degree ramp,
friction U[0.1, 0.9], initial speed U[0, 2] m/s down-slope, 120 steps, 10
iterations; avbd.scenes.ramp).

Loss: L = sum over the boxes of |x_i(T) - target_i|^2 with the seeded targets
of scenes.ramp. For each box i and each step size h in H, the oracle (colour
schedule, index claim order, the schedule the GPU's exact precision runs)
evaluates L with box i's friction at mu_i +- h and, separately, with its
initial velocity moved by +- h along the slope, and records
(L+ - L-) / (2h). A box whose differences agree across the step sizes is in
one contact regime over that range (it keeps sliding, or keeps sticking, and
meets the same neighbours); the test compares the adjoint with those.

Run from the repository root (after `make -C oracle`):
    python tests/golden/make_ramp_fd.py
"""
import json
import multiprocessing as mp
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "avbd-sandbox_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle  # noqa: E402
from avbd import scenes  # noqa: E402

H = (1e-3, 1e-4, 1e-5, 1e-6)
SCENE, PARAMS, META = scenes.ramp(64)
TARGET = np.array(META["targets"])
SLOPE = np.array(META["slope"])


def loss(scene):
    o = pyoracle.Oracle(scene, PARAMS, schedule=pyoracle.SCHED_COLOR)
    o.step(META["steps"])
    x = o.state()[2:, :3]
    return float(((x - TARGET[2:]) ** 2).sum())


def perturbed(i, what, dh):
    sc = dict(SCENE, bodies=[dict(b) for b in SCENE["bodies"]])
    b = sc["bodies"][i]
    if what == "mu":
        b["friction"] = b["friction"] + dh
    else:
        b["vel"] = list(np.array(b["vel"]) + dh * SLOPE)
    return sc


def fd(job):
    i, what, h = job
    return (loss(perturbed(i, what, h)) - loss(perturbed(i, what, -h))) / (2 * h)


def main():
    jobs = [(i, w, h) for i in range(2, len(SCENE["bodies"])) for w in ("mu", "v") for h in H]
    with mp.Pool(min(8, os.cpu_count() or 1)) as pool:
        vals = pool.map(fd, jobs)
    res = {}
    for (i, w, h), v in zip(jobs, vals):
        res.setdefault(str(i), {}).setdefault(w, {})[repr(h)] = v
    doc = dict(config="ramp", steps=META["steps"], order="index", loss=loss(SCENE), h=list(H),
               targets=META["targets"], slope=META["slope"], fd=res)
    out = os.path.join(ROOT, "tests", "golden", "configs", "ramp_fd.json")
    with open(out, "w") as f:
        json.dump(doc, f, indent=0)
    print(out, doc["loss"])


if __name__ == "__main__":
    main()
