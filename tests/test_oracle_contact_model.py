"""The oracle's GPU contact model against the reference's own checks of it.

The reference GPU backend has a contact law its CPU engine lacks: separate
static and dynamic friction, a restitution target for new impacts, a
speculative contact skin (contactOffset) with an independent rest distance
(restOffset), and a geometric fallback for warm starts when feature keys
change (shaders.js:1105-1367, 1777-1850, 2229-2264; rigid.js:144-214;
solver.js:142-162). The oracle restates it under params contactModel=1. The
reference holds no trajectories of it (the WebGPU kernels cannot run here);
it pins the law by behavioural checks, test/gpu_headless.mjs:1869-2151, and
the scenes, steps and thresholds below are those checks'. Closed forms pin
the rest (the parts a plausible mistake would break): the restitution target
e (|v| dt - gap) and the gaps it is taken from, the free flight when the skin
is 0, the equilibrium overlap set by the rest offsets, Coulomb deceleration
mu_d g of a sliding box and acceleration g (sin - mu_d cos) down a 30-degree
incline, and a hold on it when mu_s > tan 30.
"""
import math

import numpy as np
import pytest

import pyoracle

GPU = dict(contactModel=1)


def box(size, density, friction, pos, vel=(0, 0, 0), **material):
    return dict(dict(shape="box", size=list(size), density=density, friction=friction, pos=list(pos),
                     vel=list(vel)), **material)


def world(bodies, **params):
    return pyoracle.Oracle(dict(bodies=bodies, joints=[]), dict(GPU, **params))


def manifolds(o):
    kab, _ = o.forces()
    return [(int(a), int(b), o.manifold(i)) for i, (k, a, b, _) in enumerate(kab) if k == 0]


def one_step_impact(velocity, restitution, contact_offset):
    """gpu_headless.mjs:1895-1923: a unit box 6 mm above the ground, one step
    without gravity."""
    o = world([box((20, 20, 2), 0, 0.5, (0, 0, -1), restitution=restitution),
               box((1, 1, 1), 1, 0.5, (0, 0, 0.506), (0, 0, velocity), restitution=restitution)],
              gravity=0.0, iterations=10, contactOffset=contact_offset, restOffset=0.0,
              restitutionThreshold=0.25)
    o.step(1)
    st = o.state()
    ms = manifolds(o)
    return dict(z=st[1, 2], vz=st[1, 9], contacts=sum(len(m["features"]) for _, _, m in ms),
                maxBias=max([0.0] + [float(x) for _, _, m in ms for x in m["bias"]]), manifolds=ms)


def staged_impact(restitution):
    """gpu_headless.mjs:1937-1977: the box enters the 20 mm skin one step
    before it crosses the -10 mm rest distance."""
    o = world([box((20, 20, 2), 0, 0.5, (0, 0, -1), restitution=restitution),
               box((1, 1, 1), 1, 0.5, (0, 0, 0.518), (0, 0, -1), restitution=restitution)],
              gravity=0.0, iterations=10, contactOffset=0.01, restOffset=-0.005, restitutionThreshold=0.25)
    out = {}
    for s in (1, 2):
        o.step(1)
        ms = manifolds(o)
        out[s] = dict(bias=max([0.0] + [float(x) for _, _, m in ms for x in m["bias"]]),
                      C0=[m["contacts"][:, 6].copy() for _, _, m in ms], z=o.state()[1, 2], vz=o.state()[1, 9])
    return out


def test_restitution_is_a_one_step_impact_target():
    # gpu_headless.mjs:2058-2071
    inelastic = one_step_impact(-3.0, 0.0, 0.005)
    elastic = one_step_impact(-3.0, 0.8, 0.005)
    assert elastic["maxBias"] > 0.02 and elastic["vz"] > inelastic["vz"] + 1, (elastic, inelastic)
    # closed form: the skin is 2 x 5 mm, the rest distance 0, so every corner
    # contact starts 6 mm out (C0.x, the gap to the rest surface) and closes
    # |v| dt = 50 mm this step; the target is e (|v| dt - gap)
    _, _, m = elastic["manifolds"][0]
    gap = m["contacts"][:, 6]
    assert np.allclose(gap, 0.006, atol=1e-12)
    assert m["contact_distance"] == pytest.approx(0.01, abs=1e-15)
    expect = 0.8 * (3.0 / 60.0 - gap)
    assert np.allclose(m["bias"], expect, rtol=1e-12), (m["bias"], expect)
    assert inelastic["maxBias"] == 0.0


def test_restitution_stays_armed_across_a_speculative_only_step():
    # gpu_headless.mjs:2073-2084
    ine, ela = staged_impact(0.0), staged_impact(0.8)
    assert ela[1]["bias"] == 0.0 and ela[2]["bias"] > 0.002 and ela[2]["vz"] > ine[2]["vz"] + 0.1, (ela, ine)
    # step 1: 18 mm inside a 20 mm skin, C0.x = 18 mm + 10 mm to the rest
    # surface, more than the 16.7 mm closed this step; step 2 crosses it, and
    # the contact, outside rest before and unbiased, is still armed
    assert np.allclose(ela[1]["C0"][0], 0.028, atol=1e-12)
    gap2 = ela[2]["C0"][0]
    assert np.all(gap2 > 0) and np.all(gap2 < 1.0 / 60.0)


def test_speculative_contacts_and_the_skin():
    # gpu_headless.mjs:2125-2137
    discrete = one_step_impact(-1.0, 0.0, 0.0)
    spec = one_step_impact(-1.0, 0.0, 0.005)
    assert discrete["contacts"] == 0 and spec["contacts"] > 0
    assert spec["z"] > discrete["z"] + 0.005 and spec["z"] > 0.495, (discrete["z"], spec["z"])
    # without contacts the box moves freely by v dt
    assert discrete["z"] == pytest.approx(0.506 - 1.0 / 60.0, abs=1e-12)


def test_rest_offset_sets_the_equilibrium_independently_of_the_skin():
    # gpu_headless.mjs:2139-2151 (runOnGpu: default solver, 10 iterations)
    z = {}
    for rest in (-0.005, 0.0):
        o = world([box((60, 60, 2), 0, 0.5, (0, 0, -1)), box((1, 1, 1), 1, 0.5, (0, 0, 2))],
                  contactOffset=0.005, restOffset=rest)
        o.step(150)
        z[rest] = o.state()[1, 2]
    assert z[0.0] > z[-0.005] + 0.006 and abs(z[0.0] - 0.5) < 0.004, z


def test_static_and_dynamic_friction():
    # gpu_headless.mjs:1980-2055: after 90 settling steps two boxes are set
    # sliding at 10 and 200 m/s for one single-iteration step
    mat = dict(staticFriction=0.8, dynamicFriction=0.05)
    o = world([box((60, 60, 2), 0, 0.5, (0, 0, -1), **mat), box((1, 1, 1), 1, 0.8, (-2, 0, 0.48), **mat),
               box((1, 1, 1), 1, 0.8, (2, 0, 0.48), **mat)])
    o.step(90)
    o.set_params(iterations=1)
    o.set_body(1, vel=[10.0, 0.0, 0.0])
    o.set_body(2, vel=[200.0, 0.0, 0.0])
    o.step(1)
    ms = manifolds(o)

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


def test_materials_are_validated_as_set_material_does():
    # rigid.js:144-214
    for bad in (dict(staticFriction=-1.0), dict(staticFriction=0.2, dynamicFriction=0.3), dict(restitution=1.5),
                dict(contactOffset=-0.01), dict(contactOffset=0.01, restOffset=0.02),
                dict(staticFriction=math.inf)):
        with pytest.raises(ValueError):
            world([box((1, 1, 1), 1, 0.5, (0, 0, 0), **bad)])


def test_default_model_is_unchanged_by_materials():
    # the CPU engine's law reads friction alone: materials set on the bodies
    # leave its trajectory as it is, bit for bit
    base = [box((20, 20, 2), 0, 0.5, (0, 0, -1)), box((1, 1, 1), 1, 0.5, (0.1, 0, 0.6), (1, 0, 0))]
    mat = [dict(b, staticFriction=0.9, dynamicFriction=0.1, restitution=0.5, contactOffset=0.02) for b in base]
    a = pyoracle.Oracle(dict(bodies=base, joints=[]))
    b = pyoracle.Oracle(dict(bodies=mat, joints=[]))
    a.step(30)
    b.step(30)
    assert np.array_equal(a.state(), b.state())


def test_geometric_persistence_warm_starts_contacts_whose_keys_change():
    # shaders.js:1174-1319: in a falling stack of equal cubes the separating
    # axis of a pair flips between its two faces now and then, which renames
    # every contact of the pair; the geometric match carries their k and
    # lambda across (the CPU engine's feature match drops them)
    from avbd import scenes
    scene, params, _ = scenes.stack10()
    o = pyoracle.Oracle(scene, dict(params, contactModel=1, colorOrder=1), color_order=1)
    o.step(60)
    assert o.geometric_matches() > 0
    assert np.isfinite(o.state()).all()
    cpu = pyoracle.Oracle(scene, dict(params, colorOrder=1), color_order=1)
    cpu.step(60)
    assert cpu.geometric_matches() == 0


def test_rest_distance_is_the_equilibrium_overlap():
    # a resting unit box sits at 0.5 + (rest offset of the box + of the
    # ground), to the solver's convergence (1 mm)
    for rest in (-0.005, 0.0, 0.002):
        o = world([box((60, 60, 2), 0, 0.5, (0, 0, -1)), box((1, 1, 1), 1, 0.5, (0, 0, 0.6))],
                  contactOffset=0.005, restOffset=rest)
        o.step(120)
        assert abs(o.state()[1, 2] - (0.5 + 2 * rest)) < 1e-3, (rest, o.state()[1, 2])


@pytest.mark.parametrize("mu_d", [0.2, 0.4])
def test_sliding_box_decelerates_at_mu_d_g(mu_d):
    # Coulomb: a box sliding on flat ground loses mu_d g of speed per second
    mat = dict(staticFriction=0.6, dynamicFriction=mu_d)
    o = world([box((60, 60, 2), 0, 0.5, (0, 0, -1), **mat), box((1, 1, 1), 1, 0.5, (0, 0, 0.495), **mat)],
              gravity=-10.0)
    o.step(60)
    o.set_body(1, vel=[5.0, 0.0, 0.0])
    v = []
    for _ in range(20):
        o.step(1)
        v.append(o.state()[1, 7])
    dec = -np.diff(v)[4:] * 60.0
    assert np.all(np.abs(dec - mu_d * 10.0) < 0.05 * mu_d * 10.0), dec


def incline(mu_s, mu_d, steps):
    """A unit box resting on a 30-degree incline; the box's centre per step."""
    th = math.radians(30.0)
    q = [0.0, math.sin(-th / 2), 0.0, math.cos(-th / 2)]
    n = np.array([-math.sin(th), 0.0, math.cos(th)])
    mat = dict(staticFriction=mu_s, dynamicFriction=mu_d)
    o = world([dict(box((40, 40, 2), 0, 0.5, (0, 0, 0), **mat), quat=q),
               dict(box((1, 1, 1), 1, 0.5, list(n * 1.49), **mat), quat=q)], gravity=-10.0)
    xs = []
    for _ in range(steps):
        o.step(1)
        xs.append(o.state()[1, :3].copy())
    return np.array(xs)


def test_incline_static_friction_holds_and_dynamic_friction_slides():
    th = math.radians(30.0)
    # mu_s = mu_d = 0.8 > tan 30: the box stays where it settled
    held = incline(0.8, 0.8, 90)
    assert np.linalg.norm(held[-1] - held[10]) < 2e-3
    # mu_s = 0.5 < tan 30: it slides, accelerating at g (sin - mu_d cos)
    xs = incline(0.5, 0.3, 90)
    v = np.linalg.norm(np.diff(xs, axis=0), axis=1) * 60.0
    acc = np.diff(v[40:80]).mean() * 60.0
    expect = 10.0 * (math.sin(th) - 0.3 * math.cos(th))
    assert abs(acc - expect) < 0.02 * expect, (acc, expect)
