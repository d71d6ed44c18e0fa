"""The oracle against mathematics: closed forms, a library solver, brute force.

These pin what the reference-generated golden vectors leave open, and would
catch a plausible mistake in a term the fixtures exercise little (a dropped
term, a sign, a transposed operand). References: test/analytic.mjs (checks
1-3, 6), src/math/maths.js (solve6), src/physics/broadphase.js and
coloring.js.
"""
import math

import numpy as np
import pytest

import pyoracle
from avbd import scenes


def box(size, rho, mu, pos, **kw):
    return scenes.box(size, rho, mu, pos, **kw)


def test_free_fall_exact_bdf1():
    # analytic.mjs:43-56
    o = pyoracle.Oracle(dict(bodies=[box([1, 1, 1], 1, 0.5, [0, 0, 100])], joints=[]), dict(gravity=-10.0),
                        schedule=pyoracle.SCHED_REVERSE)
    n = 120
    o.step(n)
    g, dt = -10.0, 1 / 60
    z = 100 + g * dt * dt * (n * (n + 1) / 2)
    s = o.state()
    assert abs(s[0, 2] - z) / abs(z) < 1e-12
    assert abs(s[0, 9] - n * g * dt) / abs(n * g * dt) < 1e-12


def test_cuboid_and_sphere_mass_properties():
    # analytic.mjs:76-90; spheres.mjs: m = 4/3 pi r^3 rho, I = 2/5 m r^2
    sc = dict(bodies=[box([0.3, 1.7, 2.9], 7.5, 0.5, [0, 0, 0]), scenes.sphere(1.3, 2.0, 0.5, [5, 0, 0])],
              joints=[])
    m = pyoracle.Oracle(sc).mass()
    a, b, c = 0.3, 1.7, 2.9
    M = a * b * c * 7.5
    assert abs(m[0, 0] - M) < 1e-15 * M
    assert np.allclose(m[0, 1:4], [M * (b * b + c * c) / 12, M * (a * a + c * c) / 12, M * (a * a + b * b) / 12],
                       rtol=1e-15)
    assert abs(m[0, 4] - math.sqrt((a / 2) ** 2 + (b / 2) ** 2 + (c / 2) ** 2)) < 1e-15
    r = 0.65
    Ms = 4 / 3 * math.pi * r ** 3 * 2.0
    assert abs(m[1, 0] - Ms) < 1e-14 * Ms and np.allclose(m[1, 1:4], 0.4 * Ms * r * r, rtol=1e-14)


def _spd(rng, cond):
    q, _ = np.linalg.qr(rng.standard_normal((6, 6)))
    ev = np.logspace(0, math.log10(cond), 6)
    return (q * ev) @ q.T


@pytest.mark.parametrize("cond", [1e1, 1e4, 1e8, 1e10])
def test_solve6_against_lapack(cond):
    rng = np.random.default_rng(int(math.log10(cond)))
    for _ in range(20):
        A = _spd(rng, cond)
        b = rng.standard_normal(6)
        xl, xa = pyoracle.solve6(A[:3, :3], A[3:, 3:], A[3:, :3], b[:3], b[3:])
        x = np.concatenate([xl, xa])
        ref = np.linalg.solve(A, b)
        # backward error is what an LDL^T without pivoting guarantees
        assert np.linalg.norm(A @ x - b) <= 1e-12 * np.linalg.norm(A) * np.linalg.norm(x) + 1e-12
        assert np.linalg.norm(x - ref) <= cond * 1e-13 * np.linalg.norm(ref) + 1e-12


def test_solve6_reads_the_lower_triangle_only():
    rng = np.random.default_rng(5)
    A = _spd(rng, 100.0)
    b = rng.standard_normal(6)
    junk = A.copy()
    junk[np.triu_indices(6, 1)] = 1e30  # the upper triangle must not be read
    x1 = pyoracle.solve6(A[:3, :3], A[3:, 3:], A[3:, :3], b[:3], b[3:])
    x2 = pyoracle.solve6(junk[:3, :3], junk[3:, 3:], junk[3:, :3], b[:3], b[3:])
    assert np.array_equal(np.concatenate(x1), np.concatenate(x2))


def _random_scene(rng, n, spread):
    bodies = []
    for i in range(n):
        p = list(rng.uniform(-spread, spread, 3))
        if rng.random() < 0.3:
            bodies.append(scenes.sphere(rng.uniform(0.2, 1.0), 1, 0.5, p))
        else:
            q = scenes.axis_angle(*rng.standard_normal(3), rng.uniform(0, 6.28))
            bodies.append(box(list(rng.uniform(0.2, 1.0, 3)), 1, 0.5, p, quat=q))
    bodies.append(box([30, 30, 1], 0, 0.5, [0, 0, -spread - 1]))  # a large body
    return dict(bodies=bodies, joints=[])


@pytest.mark.parametrize("seed", range(4))
def test_grid_broad_phase_equals_brute_force(seed):
    # broadphase.js:236-243: the pair set and its order are the brute force's
    rng = np.random.default_rng(seed)
    sc = _random_scene(rng, 300, 4.0)
    a = pyoracle.Oracle(sc, broadphase="grid")
    b = pyoracle.Oracle(sc, broadphase="brute")
    a.record_pairs(True)
    b.record_pairs(True)
    a.step(1)
    b.step(1)
    pa, pb = a.pairs(), b.pairs()
    assert len(pa) > 100 and np.array_equal(pa, pb)
    # and the brute force is the predicate itself (solver.js:316-329)
    m = b.mass()
    pos = np.array([x["pos"] for x in sc["bodies"]])
    want = []
    for i in range(len(pos) - 1, 0, -1):
        for j in range(i - 1, -1, -1):
            d = pos[i] - pos[j]
            r = m[i, 4] + m[j, 4]
            if d @ d <= r * r and (m[i, 0] > 0 or m[j, 0] > 0):
                want.append((i, j))
    assert [tuple(p) for p in pb] == want


def _greedy(n, edges, dynamic, order):
    adj = [[] for _ in range(n)]
    for a, c in edges:
        adj[a].append(c)
        adj[c].append(a)
    col = [-1] * n
    for i in order:
        if not dynamic[i]:
            continue
        used = {col[j] for j in adj[i] if col[j] >= 0}
        c = 0
        while c in used:
            c += 1
        col[i] = c
    return col


def test_colouring_is_greedy_and_proper():
    scene, params, _ = scenes.pyramid(layers=7)
    for order in (0, 1):
        o = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR, color_order=order)
        o.step(2)
        kab, _ = o.forces()
        edges = [(a, b) for _, a, b, _ in kab if a >= 0]
        dyn = o.mass()[:, 0] > 0
        n = o.n
        if order == 0:
            visit = list(range(n))
        else:
            prio = [pyoracle.lib().orc_color_priority(i) for i in range(n)]
            visit = sorted(range(n), key=lambda i: (prio[i], i))
        want = _greedy(n, edges, dyn, visit)
        col, k = o.coloring()[0], o.coloring()[1]
        assert col.tolist() == want
        # proper: no edge between two dynamic bodies of one colour
        assert all(col[a] != col[b] for a, b in edges if dyn[a] and dyn[b])
        # the Jacobi fixed point of the GPU is the same colouring
        rank = {v: r for r, v in enumerate(visit)}
        cur = [0 if d else -1 for d in dyn]
        adj = [[] for _ in range(n)]
        for a, b in edges:
            adj[a].append(b)
            adj[b].append(a)
        for _ in range(n + 1):
            nxt = [(-1 if not dyn[i] else min(set(range(len(adj[i]) + 2)) -
                                              {cur[j] for j in adj[i] if rank[j] < rank[i] and cur[j] >= 0}))
                   for i in range(n)]
            if nxt == cur:
                break
            cur = nxt
        assert cur == want


def test_resting_box_sinks_by_the_margin():
    # analytic.mjs:185-209: a resting box settles collisionMargin into the ground
    o = pyoracle.Oracle(dict(bodies=[scenes.ground(), box([1, 1, 1], 1, 0.5, [0, 0, 0.5])], joints=[]),
                        dict(gravity=-10.0, iterations=10), schedule=pyoracle.SCHED_REVERSE)
    o.step(240)
    z = o.state()[1, 2]
    assert abs((0.5 - z) - 0.01) < 2e-4


@pytest.mark.parametrize("offset_deg,slides", [(-1.5, False), (1.5, True)])
def test_friction_threshold_on_a_slope(offset_deg, slides):
    # analytic.mjs:102-136: a box on a slope sticks below atan(sqrt(mu_a mu_b))
    # and slides above it
    mu = 0.4
    theta = math.atan(mu) + math.radians(offset_deg)
    q = scenes.axis_angle(0, 1, 0, theta)
    c, s = math.cos(theta), math.sin(theta)
    h = 0.5 + 0.5 - 0.01
    slab = box([20, 6, 1], 0, mu, [0, 0, 0], quat=q)
    cube = box([1, 1, 1], 1, mu, [s * h, 0, c * h], quat=q)
    o = pyoracle.Oracle(dict(bodies=[slab, cube], joints=[]), dict(gravity=-10.0, iterations=20),
                        schedule=pyoracle.SCHED_REVERSE)
    o.step(90)
    p = o.state()[1, :3]
    # the slab's tangent (c, 0, -s) points down the slope; the cube starts at
    # the origin of that axis
    along = p[0] * c - p[2] * s
    assert (along > 0.05) == slides, along
    if not slides:
        assert abs(along) < 0.01, along
