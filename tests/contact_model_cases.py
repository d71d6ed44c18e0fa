"""Scenes and checks of the GPU contact law, shared by the oracle's tests and
the HIP path's.

The law is the reference GPU backend's (shaders.js:1105-1367, 1777-1850,
2229-2264; rigid.js:144-214; solver.js:142-162); params contactModel=1 selects
it. The scenes and thresholds of the first five checks are the reference's
own (test/gpu_headless.mjs:1869-2151); the rest are closed forms. Each check
runs on a world made by `make(bodies, params)`, which returns an object with
step(n), state() ((n, 16): p q v w vprev), manifolds() ([(A, B, record)] with
record mu, mu_dynamic, contact_distance, bias (c,) and contacts (c, 19): rA rB
C0 k lambda C stick), set_params(**kw) and set_body(i, vel=...).
"""
import math

import numpy as np

GPU = dict(contactModel=1)


def box(size, density, friction, pos, vel=(0, 0, 0), **material):
    return dict(dict(shape="box", size=list(size), density=density, friction=friction, pos=list(pos),
                     vel=list(vel)), **material)


def scene(bodies):
    return dict(bodies=bodies, joints=[])


def max_bias(ms):
    return max([0.0] + [float(x) for _, _, m in ms for x in m["bias"]])


# ------------------------------------------------------------------ scenes
def impact_scene(velocity, restitution, contact_offset):
    """gpu_headless.mjs:1895-1923: a unit box 6 mm above the ground, no
    gravity."""
    bodies = [box((20, 20, 2), 0, 0.5, (0, 0, -1), restitution=restitution),
              box((1, 1, 1), 1, 0.5, (0, 0, 0.506), (0, 0, velocity), restitution=restitution)]
    return bodies, dict(GPU, gravity=0.0, iterations=10, contactOffset=contact_offset, restOffset=0.0,
                        restitutionThreshold=0.25)


def staged_scene(restitution):
    """gpu_headless.mjs:1937-1977: the box enters the 20 mm skin one step
    before it crosses the -10 mm rest distance."""
    bodies = [box((20, 20, 2), 0, 0.5, (0, 0, -1), restitution=restitution),
              box((1, 1, 1), 1, 0.5, (0, 0, 0.518), (0, 0, -1), restitution=restitution)]
    return bodies, dict(GPU, gravity=0.0, iterations=10, contactOffset=0.01, restOffset=-0.005,
                        restitutionThreshold=0.25)


def rest_scene(rest):
    """gpu_headless.mjs:2139-2151: a unit box dropped from 2 m."""
    return ([box((60, 60, 2), 0, 0.5, (0, 0, -1)), box((1, 1, 1), 1, 0.5, (0, 0, 2))],
            dict(GPU, contactOffset=0.005, restOffset=rest))


FRICTION_MAT = dict(staticFriction=0.8, dynamicFriction=0.05)


def friction_scene():
    """gpu_headless.mjs:1980-2003: two boxes resting on the ground."""
    m = FRICTION_MAT
    return ([box((60, 60, 2), 0, 0.5, (0, 0, -1), **m), box((1, 1, 1), 1, 0.8, (-2, 0, 0.48), **m),
             box((1, 1, 1), 1, 0.8, (2, 0, 0.48), **m)], dict(GPU))


def incline_scene(mu_s, mu_d):
    """A unit box resting on a 30-degree incline."""
    th = math.radians(30.0)
    q = [0.0, math.sin(-th / 2), 0.0, math.cos(-th / 2)]
    n = np.array([-math.sin(th), 0.0, math.cos(th)])
    mat = dict(staticFriction=mu_s, dynamicFriction=mu_d)
    return ([dict(box((40, 40, 2), 0, 0.5, (0, 0, 0), **mat), quat=q),
             dict(box((1, 1, 1), 1, 0.5, list(n * 1.49), **mat), quat=q)], dict(GPU, gravity=-10.0))


def sliding_scene(mu_d):
    mat = dict(staticFriction=0.6, dynamicFriction=mu_d)
    return ([box((60, 60, 2), 0, 0.5, (0, 0, -1), **mat), box((1, 1, 1), 1, 0.5, (0, 0, 0.495), **mat)],
            dict(GPU, gravity=-10.0))


def mixed_scene(n=48, seed=7):
    """Boxes and spheres of random material (restitution, static and dynamic
    friction, some with their own offsets) dropped into a walled pit: every
    branch of the law at once (speculative contacts, impacts, sticking and
    sliding, sphere pairs, geometric matches)."""
    rng = np.random.default_rng(seed)
    bodies = [box((12, 12, 1), 0, 0.6, (0, 0, -0.5)),
              box((1, 12, 4), 0, 0.6, (-5.5, 0, 2)), box((1, 12, 4), 0, 0.6, (5.5, 0, 2)),
              box((12, 1, 4), 0, 0.6, (0, -5.5, 2)), box((12, 1, 4), 0, 0.6, (0, 5.5, 2))]
    for i in range(n):
        sf = float(rng.uniform(0.2, 0.9))
        mat = dict(staticFriction=sf, dynamicFriction=float(sf * rng.uniform(0.3, 1.0)),
                   restitution=float(rng.choice([0.0, 0.3, 0.8])))
        if i % 5 == 0:
            mat["contactOffset"] = 0.01
        if i % 7 == 0:
            mat["restOffset"] = 0.0
        pos = [float(rng.uniform(-3.5, 3.5)), float(rng.uniform(-3.5, 3.5)), 1.0 + 0.9 * i / 4]
        vel = [float(rng.uniform(-1, 1)), float(rng.uniform(-1, 1)), float(rng.uniform(-3, 0))]
        if i % 3 == 0:
            d = float(rng.uniform(0.5, 0.9))
            bodies.append(dict(shape="sphere", size=[d, d, d], density=1.0, friction=0.5, pos=pos, vel=vel, **mat))
        else:
            s = [float(x) for x in rng.uniform(0.5, 1.0, 3)]
            bodies.append(dict(box(s, 1.0, 0.5, pos, vel, **mat)))
    return bodies, dict(GPU, gravity=-10.0, colorOrder=1)


# ------------------------------------------------------------------ checks
def one_step_impact(make, velocity, restitution, contact_offset):
    w = make(*impact_scene(velocity, restitution, contact_offset))
    w.step(1)
    st, ms = w.state(), w.manifolds()
    return dict(z=st[1, 2], vz=st[1, 9], contacts=sum(len(m["bias"]) for _, _, m in ms), maxBias=max_bias(ms),
                manifolds=ms)


def check_restitution_target(make, exact_gaps=True):
    # gpu_headless.mjs:2058-2071
    inelastic = one_step_impact(make, -3.0, 0.0, 0.005)
    elastic = one_step_impact(make, -3.0, 0.8, 0.005)
    assert elastic["maxBias"] > 0.02 and elastic["vz"] > inelastic["vz"] + 1, (elastic, inelastic)
    assert inelastic["maxBias"] == 0.0
    # closed form: every corner contact starts 6 mm from the rest surface
    # (skin 2 x 5 mm, rest distance 0) and closes |v| dt = 50 mm this step
    _, _, m = elastic["manifolds"][0]
    gap = m["contacts"][:, 6]
    tol = 1e-12 if exact_gaps else 1e-6
    assert np.allclose(gap, 0.006, atol=tol), gap
    assert abs(m["contact_distance"] - 0.01) < tol
    assert np.allclose(m["bias"], 0.8 * (3.0 / 60.0 - gap), rtol=tol, atol=tol), m["bias"]


def check_staged_restitution(make):
    # gpu_headless.mjs:2073-2084
    out = {}
    for e in (0.0, 0.8):
        w = make(*staged_scene(e))
        rec = []
        for _ in range(2):
            w.step(1)
            ms = w.manifolds()
            rec.append(dict(bias=max_bias(ms), C0=[m["contacts"][:, 6].copy() for _, _, m in ms],
                            vz=w.state()[1, 9]))
        out[e] = rec
    ine, ela = out[0.0], out[0.8]
    assert ela[0]["bias"] == 0.0 and ela[1]["bias"] > 0.002 and ela[1]["vz"] > ine[1]["vz"] + 0.1, out
    # step 1: 18 mm into a 20 mm skin, 28 mm from the -10 mm rest surface,
    # more than the 16.7 mm closed in a step; step 2 crosses it
    assert np.allclose(ela[0]["C0"][0], 0.028, atol=1e-6)
    assert np.all(ela[1]["C0"][0] > 0) and np.all(ela[1]["C0"][0] < 1.0 / 60.0)


def check_speculative_skin(make):
    # gpu_headless.mjs:2125-2137
    discrete = one_step_impact(make, -1.0, 0.0, 0.0)
    spec = one_step_impact(make, -1.0, 0.0, 0.005)
    assert discrete["contacts"] == 0 and spec["contacts"] > 0
    assert spec["z"] > discrete["z"] + 0.005 and spec["z"] > 0.495, (discrete["z"], spec["z"])
    assert abs(discrete["z"] - (0.506 - 1.0 / 60.0)) < 1e-12


def check_rest_offset_equilibrium(make):
    # gpu_headless.mjs:2139-2151
    z = {}
    for rest in (-0.005, 0.0):
        w = make(*rest_scene(rest))
        w.step(150)
        z[rest] = w.state()[1, 2]
    assert z[0.0] > z[-0.005] + 0.006 and abs(z[0.0] - 0.5) < 0.004, z
    # closed form: the resting box overlaps the ground by the two rest offsets
    for rest, zr in z.items():
        assert abs(zr - (0.5 + 2 * rest)) < 1e-3, z


def check_static_and_dynamic_friction(make):
    # gpu_headless.mjs:1980-2055
    w = make(*friction_scene())
    w.step(90)
    w.set_params(iterations=1)
    w.set_body(1, vel=[10.0, 0.0, 0.0])
    w.set_body(2, vel=[200.0, 0.0, 0.0])
    w.step(1)
    ms = w.manifolds()

    def ratio(body):
        n = t1 = t2 = 0.0
        stick = contacts = 0
        for a, b, m in ms:
            if body not in (a, b):
                continue
            for c in m["contacts"]:
                n += abs(c[12]); t1 += c[13]; t2 += c[14]
                stick += int(c[18] > 0.5); contacts += 1
        return dict(contacts=contacts, sticking=stick, ratio=math.hypot(t1, t2) / max(n, 1e-12))

    heads = [m for a, b, m in ms if 1 in (a, b) or 2 in (a, b)]
    assert heads and all(abs(m["mu"] - 0.8) < 1e-6 and abs(m["mu_dynamic"] - 0.05) < 1e-6 for m in heads)
    slow, fast = ratio(1), ratio(2)
    assert slow["contacts"] > 0 and slow["sticking"] == slow["contacts"], slow
    assert 0.05 * 1.2 < slow["ratio"] <= 0.8 * 1.05, slow
    assert fast["contacts"] > 0 and fast["sticking"] == 0 and fast["ratio"] <= 0.05 * 1.05, fast


def check_sliding_deceleration(make, mu_d):
    # Coulomb: a sliding box loses mu_d g of speed per second
    w = make(*sliding_scene(mu_d))
    w.step(60)
    w.set_body(1, vel=[5.0, 0.0, 0.0])
    v = []
    for _ in range(20):
        w.step(1)
        v.append(w.state()[1, 7])
    dec = -np.diff(v)[4:] * 60.0
    assert np.all(np.abs(dec - mu_d * 10.0) < 0.05 * mu_d * 10.0), dec


def check_incline(make):
    th = math.radians(30.0)

    def run(mu_s, mu_d, steps):
        w = make(*incline_scene(mu_s, mu_d))
        xs = []
        for _ in range(steps):
            w.step(1)
            xs.append(w.state()[1, :3].copy())
        return np.array(xs)

    # mu_s = mu_d = 0.8 > tan 30: the box stays where it settled
    held = run(0.8, 0.8, 90)
    assert np.linalg.norm(held[-1] - held[10]) < 2e-3
    # mu_s = 0.5 < tan 30: it slides at g (sin - mu_d cos)
    xs = run(0.5, 0.3, 90)
    v = np.linalg.norm(np.diff(xs, axis=0), axis=1) * 60.0
    acc = np.diff(v[40:80]).mean() * 60.0
    expect = 10.0 * (math.sin(th) - 0.3 * math.cos(th))
    assert abs(acc - expect) < 0.02 * expect, (acc, expect)
