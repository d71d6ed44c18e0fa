"""The oracle's Solver.removeBodies (solver.js:261-300), pinned by what the
reference's semantics fix: removing a body that touches nothing is the same
as never having had it (the survivors keep their order, forces keep theirs and
the bodies they name are renumbered), and a removal keeps every surviving
force's warm-start state untouched while dropping exactly the forces that
touch a removed body."""
import numpy as np

import avbd
import pyoracle

STEPS = 20


def stack_scene(extra_at=None):
    scene, params, _ = avbd.scenes.stack10()
    bodies = scene["bodies"]
    if extra_at is not None:
        # a static cube far from everything, at index extra_at
        bodies = bodies[:extra_at] + [avbd.scenes.box([1, 1, 1], 0, 0.5, [500.0, 0, 50.0])] + bodies[extra_at:]
    return dict(bodies=bodies, joints=[]), params


def test_removing_an_isolated_body_is_never_having_it():
    sa, params = stack_scene(extra_at=4)
    sb, _ = stack_scene()
    a = pyoracle.Oracle(sa, params, schedule=pyoracle.SCHED_COLOR)
    b = pyoracle.Oracle(sb, params, schedule=pyoracle.SCHED_COLOR)
    a.step(STEPS)
    doomed = np.zeros(a.n, bool)
    doomed[4] = True
    assert a.remove_bodies(doomed) == 1
    b.step(STEPS)
    assert np.array_equal(a.state(), b.state())
    a.step(STEPS)
    b.step(STEPS)
    assert np.array_equal(a.state(), b.state())
    fa, fb = a.forces(), b.forces()
    assert np.array_equal(fa[0], fb[0])
    for i in range(len(fa[0])):
        ma, mb = a.manifold(i), b.manifold(i)
        assert np.array_equal(ma["contacts"], mb["contacts"])


def test_removal_keeps_surviving_forces_and_drops_the_rest():
    scene, params, _ = avbd.scenes.pyramid(layers=5, speed=0.0)
    o = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR)
    o.step(15)
    kab, seq = o.forces()
    before = {(int(a), int(b)): o.manifold(i) for i, (k, a, b, _) in enumerate(kab) if k == 0}
    doomed = np.zeros(o.n, bool)
    doomed[[3, 7, 8]] = True
    remap = np.cumsum(~doomed) - 1
    assert o.remove_bodies(doomed) == 3
    kab2, seq2 = o.forces()
    want = [(remap[a], remap[b]) for (k, a, b, _) in kab if not doomed[a] and not doomed[b]]
    assert [(int(a), int(b)) for (_, a, b, _) in kab2] == want
    # survivors keep their creation order
    assert np.all(np.diff(seq2) > 0)
    inv = {(int(remap[a]), int(remap[b])): (a, b) for (a, b) in before if not doomed[a] and not doomed[b]}
    for i, (_, a, b, _) in enumerate(kab2):
        old = before[inv[(int(a), int(b))]]
        new = o.manifold(i)
        assert np.array_equal(old["contacts"], new["contacts"]) and np.array_equal(old["features"], new["features"])
    # appended bodies and further steps work on the compacted world
    ids = o.add_bodies([avbd.scenes.box([1, 1, 1], 1.0, 0.5, [0.0, 0.0, 12.0])])
    assert ids == [o.n - 1]
    o.step(5)
    assert np.isfinite(o.state()).all()
