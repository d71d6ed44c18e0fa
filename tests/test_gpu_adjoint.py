"""Adjoint of the step against f64 central finite differences of the oracle.

The scene is the sliding regime of the ramp benchmark: boxes side by side on a
30-degree ramp with friction below tan 30 deg, so every contact keeps sliding
and the trajectory is a smooth function of each box's friction and initial
speed (the differences agree to 7 digits for h from 1e-4 to 1e-7). The GPU
forward runs in the exact precision, bit-identical to the oracle, so the two
sides differentiate the same computed trajectory. Tolerance: 1e-4 relative to
the largest gradient of the set (the adjoint holds the contact geometry
piecewise constant; in this regime that is exact up to rounding).
"""
import copy

import numpy as np
import pytest

import avbd
import pyoracle
from avbd import AvbdError

pytestmark = pytest.mark.gpu

STEPS = 60


def final_positions(scene, params):
    o = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR)
    o.step(STEPS)
    return o.state()[:, :3]


def per_body_loss(x, target):
    return ((x - target) ** 2).sum(axis=1)


# adjoint_small: 0 reverses each step in one workgroup (small scenes), -1 with
# one kernel per phase (the path of large scenes)
@pytest.mark.parametrize("small", [0, -1], ids=["one-workgroup", "per-phase"])
@pytest.mark.parametrize("precision", [avbd.PREC_EXACT, avbd.PREC_F64], ids=["exact", "f64"])
def test_adjoint_matches_finite_differences(precision, small):
    scene, params, meta = avbd.scenes.ramp_lanes(n=8)
    slope = np.array(meta["slope"])
    n = len(scene["bodies"])
    target = np.array([b["pos"] for b in scene["bodies"]]) + np.array([2.0, 0.0, -1.0])
    gpu = avbd.HipBackend(precision=precision, adjoint_small=small)
    gpu.pack(scene, params)
    gpu.adjoint_forward(STEPS, 10)
    xg = gpu.read_state()["pos"]
    if precision == avbd.PREC_EXACT:
        assert np.array_equal(xg, final_positions(scene, params))
    dL = 2 * (xg - target)
    dL[:2] = 0  # static ground and ramp
    gmu, gv0 = gpu.adjoint_backward(dL)
    # the backward pass leaves the context at the forward's final state
    assert np.array_equal(gpu.read_state()["pos"], xg)

    h = 1e-6
    boxes = range(2, n)
    fd_mu, fd_v = [], []
    for i in boxes:
        sp, sm = copy.deepcopy(scene), copy.deepcopy(scene)
        sp["bodies"][i]["friction"] += h
        sm["bodies"][i]["friction"] -= h
        fd_mu.append((per_body_loss(final_positions(sp, params), target)[i] -
                      per_body_loss(final_positions(sm, params), target)[i]) / (2 * h))
        sp, sm = copy.deepcopy(scene), copy.deepcopy(scene)
        sp["bodies"][i]["vel"] = list(np.array(scene["bodies"][i]["vel"]) + h * slope)
        sm["bodies"][i]["vel"] = list(np.array(scene["bodies"][i]["vel"]) - h * slope)
        fd_v.append((per_body_loss(final_positions(sp, params), target)[i] -
                     per_body_loss(final_positions(sm, params), target)[i]) / (2 * h))
    fd_mu, fd_v = np.array(fd_mu), np.array(fd_v)
    ad_mu = gmu[2:]
    ad_v = gv0[2:] @ slope
    print("dL/dmu  adjoint", np.round(ad_mu, 6), "\n        fd     ", np.round(fd_mu, 6))
    print("dL/dv0  adjoint", np.round(ad_v, 6), "\n        fd     ", np.round(fd_v, 6))
    assert np.abs(ad_mu - fd_mu).max() <= 1e-4 * np.abs(fd_mu).max()
    assert np.abs(ad_v - fd_v).max() <= 1e-4 * np.abs(fd_v).max()


def test_adjoint_forward_equals_stepping_and_rejects_unsupported():
    scene, params, _ = avbd.scenes.ramp_lanes(n=4)
    a = avbd.HipBackend(precision=avbd.PREC_EXACT)
    a.pack(scene, params)
    a.adjoint_forward(25, 7)
    b = avbd.HipBackend(precision=avbd.PREC_EXACT)
    b.pack(scene, params)
    b.step(25)
    assert np.array_equal(a.read_state()["pos"], b.read_state()["pos"])
    m = avbd.HipBackend(precision=avbd.PREC_MIXED)
    m.pack(scene, params)
    with pytest.raises(AvbdError) as e:
        m.adjoint_forward(5, 5)
    assert e.value.code == -4
    with pytest.raises(AvbdError) as e:
        b.adjoint_backward(np.zeros((b.n, 3)))
    assert e.value.code == -4  # no adjoint forward recorded


def test_adjoint_paths_agree_on_colliding_boxes():
    # the full ramp benchmark (boxes run into each other): both reverse paths
    # give the same gradients up to the order of atomic additions (one
    # workgroup with barriers between phases, or a kernel per phase)
    scene, params, meta = avbd.scenes.ramp(16)
    grads = []
    for small in (0, -1):
        g = avbd.HipBackend(precision=avbd.PREC_F64, adjoint_small=small)
        g.pack(scene, params)
        g.adjoint_forward(60, 10)
        x = g.read_state()["pos"]
        dL = 2 * x
        dL[:2] = 0
        grads.append(g.adjoint_backward(dL))
    (m0, v0), (m1, v1) = grads
    assert np.isfinite(m0).all() and np.isfinite(v0).all()
    assert np.abs(m0 - m1).max() <= 1e-9 * max(1.0, np.abs(m0).max())
    assert np.abs(v0 - v1).max() <= 1e-9 * max(1.0, np.abs(v0).max())
    # static bodies never move (solver.js:396-400): their velocity is inert
    assert not v0[:2].any() and not v1[:2].any()


@pytest.mark.parametrize("small", [0, -1], ids=["one-workgroup", "per-phase"])
def test_resident_tape_equals_segment_reruns(small):
    # the same forward recorded as it runs (tape kept in device memory) or
    # re-run segment by segment from the checkpoints in the backward: the
    # reverse reads the same tape, so the gradients are the same bits
    scene, params, meta = avbd.scenes.ramp(16)
    grads = []
    for tape_mb in (0, -1):
        g = avbd.HipBackend(precision=avbd.PREC_F64, adjoint_small=small, adjoint_tape_mb=tape_mb)
        g.pack(scene, params)
        g.adjoint_forward(60, 10)
        x = g.read_state()["pos"]
        dL = 2 * x
        dL[:2] = 0
        grads.append(g.adjoint_backward(dL))
        assert np.array_equal(g.read_state()["pos"], x)
    (m0, v0), (m1, v1) = grads
    assert np.array_equal(m0, m1) and np.array_equal(v0, v1)


def test_tape_buffers_reused_across_packs():
    # a context keeps its tape's buffers across packs (an optimisation loop
    # re-packs and records again): gradients equal those of a fresh context,
    # also when the second recording is shorter than the first
    scene, params, meta = avbd.scenes.ramp(16)

    def grads(g, steps):
        g.pack(scene, params)
        g.adjoint_forward(steps, 10)
        x = g.read_state()["pos"]
        dL = 2 * x
        dL[:2] = 0
        return g.adjoint_backward(dL)

    reused = avbd.HipBackend(precision=avbd.PREC_F64)
    grads(reused, 60)
    m1, v1 = grads(reused, 40)
    m0, v0 = grads(avbd.HipBackend(precision=avbd.PREC_F64), 40)
    assert np.array_equal(m0, m1) and np.array_equal(v0, v1)


def test_recording_dropped_by_state_changes():
    # a recorded forward is reversed only while the context still holds what
    # it stepped through: parameters, state or further steps in between make
    # the backward fail before any launch, and the forward itself starts from
    # the packed scene (its v0 gradient is that of the packed velocity)
    scene, params, _ = avbd.scenes.ramp_lanes(n=4)
    g = avbd.HipBackend(precision=avbd.PREC_F64)
    dL = np.zeros((len(scene["bodies"]), 3))
    changes = {
        "set_params": lambda: g.set_params(iterations=12),
        "step": lambda: g.step(1),
        "write_state": lambda: g.write_state(**{k: v for k, v in g.read_state().items()}),
        "upload_coloring": lambda: g.upload_coloring(g.read_coloring()[0]),
    }
    for name, change in changes.items():
        g.pack(scene, params)
        g.adjoint_forward(10, 5)
        change()
        with pytest.raises(AvbdError) as e:
            g.adjoint_backward(dL)
        assert e.value.code == -4, name
    g.pack(scene, params)
    g.step(1)
    with pytest.raises(AvbdError) as e:
        g.adjoint_forward(10, 5)
    assert e.value.code == -4
    # after a fresh pack the same context records and reverses again
    g.pack(scene, params)
    g.adjoint_forward(10, 5)
    mu, v0 = g.adjoint_backward(dL)
    assert np.isfinite(mu).all() and np.isfinite(v0).all()


def test_tape_budget_counted_as_the_forward_runs():
    # boxes dropped onto the ramp: the first steps hold no contacts, so the
    # tape grows as they land. A budget the growing tape crosses part-way
    # releases it and the backward re-runs the segments; the gradients are
    # the bits of the all-re-run path
    scene, params, meta = avbd.scenes.ramp(16)
    for b in scene["bodies"][2:]:
        b["pos"] = [b["pos"][0], b["pos"][1], b["pos"][2] + 1.5]
    grads = []
    for tape_mb in (1, -1, 0):
        g = avbd.HipBackend(precision=avbd.PREC_F64, adjoint_tape_mb=tape_mb)
        g.pack(scene, params)
        g.adjoint_forward(60, 10)
        x = g.read_state()["pos"]
        dL = 2 * x
        dL[:2] = 0
        grads.append(g.adjoint_backward(dL))
    (m0, v0), (m1, v1), (m2, v2) = grads
    assert np.array_equal(m0, m1) and np.array_equal(v0, v1)
    assert np.array_equal(m0, m2) and np.array_equal(v0, v2)


def test_config5_gradients_against_finite_differences():
    # The differentiable configuration as configured: 64 boxes on the
    # 30-degree ramp, friction U[0.1, 0.9] (many above tan 30 deg, so they
    # stick), initial speed U[0, 2] m/s, 120 steps, checkpoint every 10, the
    # loss sum |x_i(T) - target_i|^2 over seeded targets. The oracle's central
    # differences (tests/golden/configs/ramp_fd.json, make_ramp_fd.py) are
    # taken at h = 1e-3 ... 1e-6; a box whose differences at h = 1e-5 and
    # 1e-6 agree is in one contact regime over that range, and its adjoint
    # gradient is compared with the difference at 1e-6. The exact precision's
    # forward is the oracle's trajectory bit for bit (the loss is compared
    # exactly), so both sides differentiate the same computation.
    import json
    import os
    doc = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "configs", "ramp_fd.json")))
    scene, params, meta = avbd.scenes.ramp(64)
    target = np.array(meta["targets"])
    slope = np.array(meta["slope"])
    gpu = avbd.HipBackend(precision=avbd.PREC_EXACT)
    gpu.pack(scene, params)
    gpu.adjoint_forward(meta["steps"], 10)
    x = gpu.read_state()["pos"]
    assert float(((x[2:] - target[2:]) ** 2).sum()) == doc["loss"]
    dL = 2 * (x - target)
    dL[:2] = 0
    gmu, gv0 = gpu.adjoint_backward(dL)
    ad = {"mu": gmu, "v": gv0 @ slope}
    report = {}
    for what in ("mu", "v"):
        fd6 = np.array([doc["fd"][str(i)][what][repr(1e-6)] for i in range(2, 66)])
        fd5 = np.array([doc["fd"][str(i)][what][repr(1e-5)] for i in range(2, 66)])
        scale = np.abs(fd6).max()
        stable = np.abs(fd5 - fd6) <= 1e-4 * np.maximum(np.abs(fd6), 1e-3 * scale)
        err = np.abs(ad[what][2:] - fd6) / np.maximum(np.abs(fd6), 1e-2 * scale)
        report[what] = dict(stable=int(stable.sum()), within_1e3=int((err[stable] <= 1e-3).sum()),
                            within_1e2=int((err[stable] <= 1e-2).sum()), within_1e1=int((err[stable] <= 1e-1).sum()),
                            worst=float(err[stable].max()), median=float(np.median(err[stable])))
    print("[config5] " + json.dumps(report), flush=True)
    # measured on MI355X: 52 of 64 friction and 58 of 64 velocity
    # differences are regime-stable; median relative error 4e-5; 43 and 51 of
    # them within 1%, 50 and 56 within 10%, the worst 27% and 29% (boxes that
    # run into a neighbour: the adjoint holds each contact's normal and anchors
    # fixed over a step, while the finite difference also moves them)
    for what in ("mu", "v"):
        r = report[what]
        assert r["stable"] >= 45, r
        assert r["median"] <= 1e-3, r
        assert r["within_1e2"] >= 0.75 * r["stable"], r
        assert r["within_1e1"] >= 0.9 * r["stable"], r
