"""The verification hooks: a caller's colouring (GpuBackend.uploadColoring)
and a caller's claim order for the greedy colouring (the device counterpart of
Solver.bodyOrder), checked against the oracle and against brute force."""
import json
import os

import numpy as np
import pytest

import avbd
import pyoracle

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_uploaded_distinct_colours_equal_sequential_order():
    # one colour per dynamic body: the sweep visits the bodies one at a time in
    # colour order, which is the oracle's explicit body-order schedule
    doc = json.load(open(os.path.join(GOLDEN, "stack10_color.json")))
    scene, params = doc["scene"], doc["params"]
    n = len(scene["bodies"])
    dyn = [i for i, b in enumerate(scene["bodies"]) if b["density"] > 0]
    order = list(reversed(dyn))[::2] + list(reversed(dyn))[1::2]
    color = np.full(n, -1, np.int32)
    for k, i in enumerate(order):
        color[i] = k
    gpu = avbd.HipBackend(precision=avbd.PREC_EXACT)
    gpu.pack(scene, params)
    gpu.upload_coloring(color)
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_ORDER)
    ref.set_schedule(pyoracle.SCHED_ORDER, order)
    for s in range(1, 31):
        gpu.step(1)
        ref.step(1)
        st = gpu.read_state()
        o = ref.state()
        assert np.array_equal(st["pos"], o[:, :3]) and np.array_equal(st["quat"], o[:, 3:7]), f"step {s}"
    assert gpu.counters()["overflow"] & 16 == 0


def test_improper_upload_falls_back_to_greedy():
    doc = json.load(open(os.path.join(GOLDEN, "stack10_color.json")))
    scene, params = doc["scene"], doc["params"]
    n = len(scene["bodies"])
    a = avbd.HipBackend(precision=avbd.PREC_EXACT)
    a.pack(scene, params)
    a.upload_coloring(np.zeros(n, np.int32))  # touching cubes in one colour
    b = avbd.HipBackend(precision=avbd.PREC_EXACT)
    b.pack(scene, params)
    a.step(10)
    b.step(10)
    assert np.array_equal(a.read_state()["pos"], b.read_state()["pos"])
    assert a.counters()["overflow"] & 16
    a.upload_coloring(None)
    with pytest.raises(avbd.AvbdError):
        a.upload_coloring(np.full(n, 300, np.int32))


@pytest.mark.parametrize("seed", [1, 2])
def test_custom_claim_order_is_greedy_in_that_order(seed):
    scene, params, _ = avbd.scenes.pyramid(layers=6)
    n = len(scene["bodies"])
    gpu = avbd.HipBackend(precision=avbd.PREC_EXACT)
    gpu.pack(scene, params)
    gpu.step(2)
    # the index order given as ranks colours exactly as the index claim order
    idx = avbd.HipBackend(precision=avbd.PREC_EXACT)
    idx.pack(scene, dict(params, colorOrder=0))
    idx.set_body_order(np.arange(n, dtype=np.int32))
    idx.step(3)
    ref = avbd.HipBackend(precision=avbd.PREC_EXACT)
    ref.pack(scene, dict(params, colorOrder=0))
    ref.step(3)
    assert np.array_equal(idx.read_coloring()[0], ref.read_coloring()[0])
    assert np.array_equal(idx.read_state()["pos"], ref.read_state()["pos"])
    # a random order: the colouring is the greedy one in that order
    rank = np.random.default_rng(seed).permutation(n).astype(np.int32)
    gpu.set_body_order(rank)
    gpu.step(1)
    col, k = gpu.read_coloring()
    kab, _ = gpu.read_forces()
    nbr = [set() for _ in range(n)]
    for kind, a, b, _ in kab:
        if a >= 0 and b >= 0:
            nbr[a].add(b)
            nbr[b].add(a)
    dyn = [i for i, b in enumerate(scene["bodies"]) if b["density"] > 0]
    want = np.full(n, -1)
    for i in sorted(dyn, key=lambda i: (rank[i], i)):
        used = {want[j] for j in nbr[i] if want[j] >= 0}
        c = 0
        while c in used:
            c += 1
        want[i] = c
    assert np.array_equal(col, want) and k == want.max() + 1


def test_read_manifolds_equals_one_by_one_reads():
    scene, params, _ = avbd.scenes.pyramid(layers=5)
    for prec in (avbd.PREC_EXACT, avbd.PREC_MIXED):
        g = avbd.HipBackend(precision=prec)
        g.pack(scene, params)
        g.step(6)
        kab, feats = g.read_forces()
        head, bm, con = g.read_manifolds()
        nj = int((kab[:, 0] != 0).sum())
        assert len(head) == len(kab) - nj and np.array_equal(head[:, :3], kab[nj:, 1:4])
        k = 0
        for t in range(len(head)):
            m = g.read_manifold(nj + t)
            assert np.array_equal(bm[t, :9], m["basis"].reshape(9)) and bm[t, 9] == m["mu"]
            nc = len(m["contacts"])
            assert np.array_equal(con[k:k + nc], m["contacts"])
            k += nc
        assert k == len(con)


def test_coloring_limits():
    # GpuBackend.colorRounds / encColors: a cap on the colours leaves none at
    # or above it and reports the possible conflicts (bit 32); a cap on the
    # relaxation rounds stops it early on a chain linked in claim order (each
    # round settles one more link); lifting both returns to the greedy
    # colouring, equal to the oracle's
    scene, params, _ = avbd.scenes.pyramid(layers=6)
    params = dict(params, colorOrder=1)
    g = avbd.HipBackend(precision=avbd.PREC_EXACT)
    g.pack(scene, params)
    g.step(3)
    _, k = g.read_coloring()
    assert k > 3
    g.set_coloring_limits(colors=3)
    g.step(1)
    col, k3 = g.read_coloring()
    assert col.max() <= 2 and k3 <= 3 and g.counters()["overflow"] & 32
    assert np.isfinite(g.read_state()["pos"]).all()
    g.set_coloring_limits(rounds=1)
    g.step(1)
    assert g.counters()["color_rounds"] <= 1
    g.set_coloring_limits()
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR, color_order=1)
    # the inertial target is readable and finite
    y, qy = g.read_inertial_target()
    assert np.isfinite(y).all() and np.abs(np.linalg.norm(qy, axis=1) - 1).max() < 1e-9


@pytest.mark.parametrize("order", [0, 1, 2], ids=["index", "priority", "rank"])
def test_one_launch_colouring_equals_the_relaxation_rounds(order):
    # the colouring in one launch (bodies in claim order, each coloured once
    # its claims-before neighbours are) is the greedy colouring the relaxation
    # rounds reach: forcing the rounds (a round limit far above the rounds
    # needed) gives the same colours at every step, as does the oracle
    scene, params, _ = avbd.scenes.pyramid(layers=10)
    params = dict(params, colorOrder=min(order, 1))
    n = len(scene["bodies"])
    rank = np.random.default_rng(7).permutation(n).astype(np.int32)
    dag = avbd.HipBackend(precision=avbd.PREC_EXACT)
    rounds = avbd.HipBackend(precision=avbd.PREC_EXACT)
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR, color_order=min(order, 1))
    for g in (dag, rounds):
        g.pack(scene, params)
        if order == 2:
            g.set_body_order(rank)
    rounds.set_coloring_limits(rounds=1 << 20)
    for s in range(1, 9):
        dag.step(1)
        rounds.step(1)
        ref.step(1)
        cd, kd = dag.read_coloring()
        cr, kr = rounds.read_coloring()
        assert kd == kr and np.array_equal(cd, cr), f"step {s}: the two colourings differ"
        if order < 2:
            co, ko = ref.last_coloring()
            assert kd == ko and np.array_equal(cd, co), f"step {s}: differs from the oracle"
        # the one launch reports its longest claims-before chain; the rounds,
        # relaxing in place from the last step's colours, often need fewer
        assert dag.counters()["color_rounds"] >= 1 and rounds.counters()["color_rounds"] >= 1
        assert dag.counters()["overflow"] == 0
    assert np.array_equal(dag.read_state()["pos"], rounds.read_state()["pos"])
