"""Whole steps replayed as graphs (avbd_config.graphs = AVBD_GRAPHS_STEP).

Such a step makes no host read: the narrow phase, the manifold compaction and
the force lists take their counts from device memory, the colouring launches
a fixed number of relaxation rounds and the sweep a fixed number of colour
launches per iteration (GpuBackend's indirect dispatch, backend.js:1496-1535).
Checked here: the same bits as steps with host reads (exact precision: the
oracle's) over runs of steps enqueued together, several steps in flight; the
caller's colouring limits; a step that needed more colours or rounds than it
launched reports it and the next graph launches more; a capacity overflow is
reported by a later step call and a repack grows the arena.
"""
import numpy as np
import pytest

import avbd
import pyoracle
from avbd import AvbdError
from contact_model_cases import mixed_scene

pytestmark = pytest.mark.gpu


def state_of(g):
    st = g.read_state()
    return np.concatenate([st["pos"], st["quat"], st["vel"], st["angvel"], st["prev_vel"]], axis=1)


def oracle_state(o):
    return o.state()[:, :16]


def run_pair(scene, params, precision, chunks, scene_solve=0):
    """The scene on a context that replays whole steps and one that reads its
    counts on the host, stepped in the given chunks; both states after each."""
    a = avbd.HipBackend(precision=precision, graphs=avbd.GRAPHS_STEP, scene_solve=scene_solve)
    b = avbd.HipBackend(precision=precision, graphs=avbd.GRAPHS_SOLVE, scene_solve=scene_solve)
    a.pack(scene, params)
    b.pack(scene, params)
    out = []
    for n in chunks:
        a.step(n)
        b.step(n)
        out.append((state_of(a), state_of(b), a.counters(), b.counters()))
    return a, b, out


CASES = {
    # name: (scene, params, chunks); chunks above 4 keep several steps in flight
    "pyramid8": (lambda: avbd.scenes.pyramid(layers=8, speed=0.0)[0], dict(gravity=-9.81, iterations=20,
                                                                            colorOrder=1), [1, 9, 10]),
    "stack10": (lambda: avbd.scenes.stack10()[0], dict(gravity=-9.81, iterations=10), [1, 12, 7]),
    "rain400": (lambda: avbd.scenes.rain(n=400, seed=3, width=8.0, height=14.0)[0],
                dict(gravity=-9.81, iterations=20, colorOrder=1), [1, 15, 14]),
    "fabric": (lambda: avbd.scenes.fabric(count=300, balls=2)[0], dict(gravity=-9.81, iterations=10), [1, 10, 9]),
    "gpu_law": (lambda: dict(bodies=mixed_scene(n=36)[0], joints=[]), mixed_scene(n=36)[1], [1, 11, 8]),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_exact_graph_steps_equal_the_oracle(name):
    make, params, chunks = CASES[name]
    scene = make()
    print(f"[step graphs] {name} {len(scene['bodies'])} bodies, chunks {chunks}", flush=True)
    a, b, out = run_pair(scene, params, avbd.PREC_EXACT, chunks)
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR, color_order=params.get("colorOrder", 0))
    for k, (sa, sb, ca, cb) in enumerate(out):
        ref.step(chunks[k])
        assert np.array_equal(sa, sb), f"chunk {k}: graph steps differ from steps with host reads"
        assert np.array_equal(sa, oracle_state(ref)), f"chunk {k}: graph steps differ from the oracle"
        for f in ("steps", "pairs", "new_pairs", "manifolds", "contacts", "colors"):
            assert ca[f] == cb[f], f"chunk {k}: {f} {ca[f]} != {cb[f]}"
        assert ca["overflow"] == 0 and cb["overflow"] == 0
    gc, gn = a.read_coloring()
    rc, rn = ref.last_coloring()
    assert gn == rn and np.array_equal(gc, rc)


def test_throughput_precision_graph_steps_track_the_host_read_steps():
    # mixed precision: the team shapes of a graph are set at its capture, so
    # its reduction trees can differ from the host-read steps'; before the
    # first feature flip (a tie broken the other way) the poses agree closely
    scene = avbd.scenes.pyramid(layers=8, speed=0.0)[0]
    params = dict(gravity=-9.81, iterations=20, colorOrder=1)
    _, _, out = run_pair(scene, params, avbd.PREC_MIXED, [1, 1])
    for sa, sb, ca, cb in out:
        assert np.abs(sa[:, :7] - sb[:, :7]).max() < 1e-4
        assert ca["overflow"] == 0


def test_one_launch_batch_graph_steps_equal_the_oracle():
    scene, params, meta = avbd.scenes.batched(num_scenes=6, seed=5)
    a, b, out = run_pair(scene, params, avbd.PREC_EXACT, [1, 9])
    for sa, sb, ca, cb in out:
        assert np.array_equal(sa, sb)
        assert ca["overflow"] == 0
    base = 0
    for part in meta["parts"]:
        ref = pyoracle.Oracle(part, params, schedule=pyoracle.SCHED_COLOR)
        ref.step(10)
        n = len(part["bodies"])
        assert np.array_equal(out[-1][0][base:base + n], oracle_state(ref))
        base += n


def test_caller_limits_are_the_graphs_launch_counts():
    # GpuBackend.colorRounds / encColors (backend.js:256-270): graph steps run
    # those rounds and colours, as steps with host reads do. The colouring
    # they leave is not proper: bodies sharing a force are swept at once (the
    # conflicts the reference tolerates), so the bits depend on timing and
    # only the limits, the report and the scale of the result are compared.
    scene = avbd.scenes.pyramid(layers=8, speed=0.0)[0]
    params = dict(gravity=-9.81, iterations=10, colorOrder=1)
    a = avbd.HipBackend(precision=avbd.PREC_EXACT, graphs=avbd.GRAPHS_STEP)
    b = avbd.HipBackend(precision=avbd.PREC_EXACT, graphs=avbd.GRAPHS_SOLVE)
    for g in (a, b):
        g.pack(scene, params)
        g.set_coloring_limits(rounds=3, colors=4)
        g.step(12)
    sa, sb = state_of(a), state_of(b)
    assert np.isfinite(sa).all() and np.abs(sa[:, :3] - sb[:, :3]).max() < 1e-3
    ca, cb = a.counters(), b.counters()
    assert ca["colors"] <= 4 and cb["colors"] <= 4
    assert ca["overflow"] & 32 and cb["overflow"] & 32  # the limits left conflicts


def lowbias32(x):
    x &= 0xffffffff
    x ^= x >> 16
    x = (x * 0x7feb352d) & 0xffffffff
    x ^= x >> 15
    x = (x * 0x846ca68b) & 0xffffffff
    x ^= x >> 16
    return x


def greedy_colouring(n, dynamic, forces):
    """The greedy colouring of a force list in the priority claim order
    (coloring.js:75-92; lowbias32 of the index, ties to the lower index,
    layout.js:50-57): each dynamic body, in that order, takes the smallest
    colour no dynamic neighbour claimed before it has."""
    nbr = [[] for _ in range(n)]
    for _, a, b, _ in forces:
        if a >= 0 and dynamic[a] and dynamic[b]:
            nbr[a].append(b)
            nbr[b].append(a)
    col = np.full(n, -1)
    for i in sorted(range(n), key=lambda i: (lowbias32(i), i)):
        if not dynamic[i]:
            continue
        used = {col[j] for j in nbr[i]}
        c = 0
        while c in used:
            c += 1
        col[i] = c
    return col


def test_too_few_colour_launches_are_reported_and_raised():
    # a pyramid pulled apart (gaps between all cubes) that falls together:
    # the first step (host reads) sees no cube touching another, one colour,
    # so the first graphs launch two; as the layers land, steps need more,
    # are capped (bit 32) and report it, and later graphs launch enough: the
    # last step's colouring is the greedy one of its forces
    scene = avbd.scenes.pyramid(layers=6, speed=0.0)[0]
    for b in scene["bodies"][1:-1]:
        b["pos"] = [b["pos"][0] * 1.1, b["pos"][1] * 1.1, b["pos"][2] * 1.6]
    params = dict(gravity=-9.81, iterations=10, colorOrder=1)
    g = avbd.HipBackend(precision=avbd.PREC_EXACT, graphs=avbd.GRAPHS_STEP)
    g.pack(scene, params)
    g.step(1)
    assert g.counters()["colors"] == 1
    g.step(120)
    c = g.counters()
    assert c["overflow"] & 32, "a step needed more colours than its graph launched"
    assert c["colors"] > 2
    forces, _ = g.read_forces()
    dynamic = np.array([b["density"] > 0 for b in scene["bodies"]])
    col, k = g.read_coloring()
    want = greedy_colouring(len(dynamic), dynamic, forces)
    assert k == want.max() + 1
    assert np.array_equal(col[dynamic], want[dynamic])
    assert np.isfinite(state_of(g)).all()


def test_capacity_overflow_is_reported_by_a_later_step():
    # the sphere lands on the pyramid after the first step: its new pairs
    # exceed a pair capacity the first step fitted in; the graph steps drop the
    # excess (bit 1 of the overflow word), a later step call fails with
    # AVBD_E_CAPACITY, and a repack grows the arena
    scene, params, meta = avbd.scenes.pyramid(layers=6, speed=0.0)
    L = meta["launch"]
    bodies = [dict(b) for b in scene["bodies"]]
    bodies[L["body"]]["pos"] = [0.0, 0.0, 7.2]
    bodies[L["body"]]["vel"] = [0.0, 0.0, -30.0]
    scene = dict(bodies=bodies, joints=[])
    probe = avbd.HipBackend(precision=avbd.PREC_EXACT)
    probe.pack(scene, params)
    probe.step(1)
    first = probe.counters()["pairs"]
    probe.step(3)
    later = probe.counters()["pairs"]
    assert later > first, "the strike adds pairs"
    g = avbd.HipBackend(precision=avbd.PREC_EXACT, graphs=avbd.GRAPHS_STEP, pair_cap=first)
    g.pack(scene, params)
    g.step(1)
    raised = False
    try:
        for _ in range(12):
            g.step(1)
    except AvbdError as e:
        assert e.code == -3
        raised = True
    assert raised
    assert g.counters()["overflow"] & 1
    assert np.isfinite(state_of(g)).all()
    g.repack(scene, list(range(len(bodies))))
    g.step(4)
    assert g.counters()["overflow"] == 0


def test_a_huge_round_limit_launches_at_most_n_plus_two_rounds():
    # a step graph holds one launch per relaxation round: a caller's limit far
    # above what any colouring needs is held to N + 2 rounds (the graph would
    # otherwise hold a million launches); the colouring is then the greedy one
    scene = avbd.scenes.pyramid(layers=4, speed=0.0)[0]
    params = dict(gravity=-9.81, iterations=10, colorOrder=1)
    g = avbd.HipBackend(precision=avbd.PREC_EXACT, graphs=avbd.GRAPHS_STEP)
    g.pack(scene, params)
    g.set_coloring_limits(rounds=1 << 20)
    g.step(6)
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR, color_order=1)
    ref.step(6)
    assert np.array_equal(state_of(g), oracle_state(ref))
    assert g.counters()["overflow"] == 0
