"""The GPU contact law on the HIP path (params contactModel=1).

Exact precision against the oracle, bit for bit, on scenes that exercise
every part of the law: speculative contacts inside the skin, the rest
distance, restitution targets (new impacts, and impacts kept armed across a
speculative-only step), the stick-selected static/dynamic cone and its dual
update, the geometric warm-start fallback, sphere pairs, per-body offsets and
the projection pass of post-stabilisation. Per step: the broad-phase pair
list, the force list with its feature keys, every body's state; and the
restitution targets and pair materials of the slots.

Then the reference's own behavioural checks of the law
(test/gpu_headless.mjs:1869-2151) and the closed forms of
tests/contact_model_cases.py, run on the throughput precision (mixed: f32
contact records, f64 state).
"""
import hashlib
import math

import numpy as np
import pytest

import avbd
import pyoracle
from contact_model_cases import (GPU, check_incline, check_rest_offset_equilibrium, check_restitution_target,
                                 check_sliding_deceleration, check_speculative_skin, check_staged_restitution,
                                 check_static_and_dynamic_friction, friction_scene, impact_scene, incline_scene,
                                 mixed_scene, rest_scene, staged_scene)

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


class GpuWorld:
    """The HIP backend behind the checks' world interface."""

    precision = avbd.PREC_MIXED

    def __init__(self, bodies, params):
        self.g = avbd.HipBackend(precision=self.precision)
        self.g.pack(dict(bodies=bodies, joints=[]), params)

    def step(self, n=1):
        self.g.step(n)

    def state(self):
        st = self.g.read_state()
        return np.concatenate([st["pos"], st["quat"], st["vel"], st["angvel"], st["prev_vel"]], axis=1)

    def manifolds(self):
        head, bm, con = self.g.read_manifolds()
        mat, bias = self.g.read_contact_model()
        out, k = [], 0
        for t in range(len(head)):
            a, b, nc = int(head[t, 0]), int(head[t, 1]), int(head[t, 2])
            c = np.zeros((nc, 19))
            c[:, :15] = con[k:k + nc, :15]
            c[:, 18] = con[k:k + nc, 15]
            out.append((a, b, dict(basis=bm[t, :9].reshape(3, 3), mu=mat[t, 0], mu_dynamic=mat[t, 1],
                                   contact_distance=mat[t, 2], contacts=c, bias=bias[k:k + nc])))
            k += nc
        return out

    def set_params(self, **kw):
        self.g.set_params(**kw)

    def set_body(self, i, vel):
        st = self.g.read_state()
        st["vel"][i] = vel
        st["prev_vel"][i] = vel
        self.g.write_state(vel=st["vel"], prev_vel=st["prev_vel"])


class ExactWorld(GpuWorld):
    precision = avbd.PREC_EXACT


# ----------------------------------------------------- exact against oracle
def stack_scene():
    from avbd import scenes
    scene, params, _ = scenes.stack10()
    return scene["bodies"], dict(params, contactModel=1, colorOrder=1)


def post_stab_scene():
    """Three boxes with restitution under post-stabilisation: the projection
    pass leaves the restitution target out."""
    b = [dict(shape="box", size=[20, 20, 2], density=0, friction=0.5, pos=[0, 0, -1], restitution=0.5)]
    for k in range(3):
        b.append(dict(shape="box", size=[1, 1, 1], density=1, friction=0.5, pos=[0.05 * k, 0, 0.6 + 1.1 * k],
                      vel=[0, 0, -2.0], restitution=0.5))
    return b, dict(GPU, postStabilize=True)


CASES = {
    # name: (bodies, params, steps, events {step: [(body, vel)]}, {step: params})
    "impact": (*impact_scene(-3.0, 0.8, 0.005), 30, {}, {}),
    "staged": (*staged_scene(0.8), 30, {}, {}),
    "rest": (*rest_scene(0.0), 150, {}, {}),
    "friction": (*friction_scene(), 91, {90: [(1, [10.0, 0, 0]), (2, [200.0, 0, 0])]}, {90: dict(iterations=1)}),
    "incline": (*incline_scene(0.5, 0.3), 90, {}, {}),
    "stack10": (*stack_scene(), 120, {}, {}),
    "mixed": (*mixed_scene(), 150, {}, {}),
    "post_stab": (*post_stab_scene(), 60, {}, {}),
}


def oracle_slots(o):
    kab, _ = o.forces()
    feats = np.zeros((len(kab), 8), np.uint32)
    mats, bias = [], []
    for i, (kind, a, b, nc) in enumerate(kab):
        if kind == 0:
            m = o.manifold(i)
            feats[i, :len(m["features"])] = m["features"]
            mats.append([m["mu"], m["mu_dynamic"], m["contact_distance"]])
            bias.extend(m["bias"])
    return kab, feats, np.array(mats).reshape(-1, 3), np.array(bias)


@pytest.mark.parametrize("name", sorted(CASES))
def test_exact_reproduces_the_oracle_under_the_gpu_law(name):
    bodies, params, steps, events, changes = CASES[name]
    print(f"[contact-law] exact {name}: {len(bodies)} bodies, {steps} steps", flush=True)
    order = int(params.get("colorOrder", 0))
    ref = pyoracle.Oracle(dict(bodies=bodies, joints=[]), params, schedule=pyoracle.SCHED_COLOR,
                          color_order=order)
    ref.record_pairs(True)
    w = ExactWorld(bodies, params)
    for s in range(1, steps + 1):
        w.step(1)
        ref.step(1)
        assert np.array_equal(w.g.read_pairs(), ref.pairs()), f"step {s}: pair lists differ"
        gk, gf = w.g.read_forces()
        rk, rf, rmat, rbias = oracle_slots(ref)
        assert np.array_equal(gk, rk), f"step {s}: force lists differ"
        assert np.array_equal(gf, rf), f"step {s}: feature keys differ"
        gst = w.state()[:, :13] + 0.0
        rst = ref.state()[:, :13] + 0.0
        if sha(gst) != sha(rst):
            d = np.abs(gst - rst)
            i, k = np.unravel_index(np.argmax(d), d.shape)
            pytest.fail(f"step {s}: state differs, max |diff| {d.max():.3e} at body {i} component {k}")
        gmat, gbias = w.g.read_contact_model()
        assert np.array_equal(gmat, rmat), f"step {s}: pair materials differ"
        assert np.array_equal(gbias, rbias), f"step {s}: restitution targets differ"
        for b, v in events.get(s, []):
            w.set_body(b, vel=v)
            ref.set_body(b, vel=v)
        if s in changes:
            w.set_params(**changes[s])
            ref.set_params(**changes[s])
    if name == "stack10":
        assert ref.geometric_matches() > 0  # the fallback was exercised
    if name == "mixed":
        assert ref.geometric_matches() > 0 and rbias.size > 0


# ------------------------------------------------ the reference's checks
def test_restitution_is_a_one_step_impact_target():
    check_restitution_target(GpuWorld, exact_gaps=False)


def test_restitution_stays_armed_across_a_speculative_only_step():
    check_staged_restitution(GpuWorld)


def test_speculative_contacts_and_the_skin():
    check_speculative_skin(GpuWorld)


def test_rest_offset_sets_the_equilibrium_independently_of_the_skin():
    check_rest_offset_equilibrium(GpuWorld)


def test_static_and_dynamic_friction():
    check_static_and_dynamic_friction(GpuWorld)


@pytest.mark.parametrize("mu_d", [0.2, 0.4])
def test_sliding_box_decelerates_at_mu_d_g(mu_d):
    check_sliding_deceleration(GpuWorld, mu_d)


def test_incline_static_friction_holds_and_dynamic_friction_slides():
    check_incline(GpuWorld)


# -------------------------------------------------------- boundary checks
def test_materials_are_validated_at_pack():
    g = avbd.HipBackend(precision=avbd.PREC_MIXED)
    base = [dict(shape="box", size=[1, 1, 1], density=1, friction=0.5, pos=[0, 0, 0])]
    for bad in (dict(staticFriction=-1.0), dict(staticFriction=0.2, dynamicFriction=0.3), dict(restitution=1.5),
                dict(contactOffset=-0.01), dict(contactOffset=0.01, restOffset=0.02),
                dict(staticFriction=math.inf)):
        with pytest.raises(avbd.AvbdError):
            g.pack(dict(bodies=[dict(base[0], **bad)], joints=[]), dict(GPU))
    with pytest.raises(avbd.AvbdError):
        g.set_params(contactModel=2)


def test_the_adjoint_refuses_the_gpu_law():
    bodies, params = rest_scene(0.0)
    g = avbd.HipBackend(precision=avbd.PREC_F64)
    g.pack(dict(bodies=bodies, joints=[]), params)
    with pytest.raises(avbd.AvbdError):
        g.adjoint_forward(2, 1)


def test_set_params_regrows_the_grid_for_a_larger_skin():
    # the pair test and the broad-phase grid follow a contact offset changed
    # after pack: two unit boxes 3.6 m apart (bounding radii 0.866) are no
    # pair at the default 5 mm skin, and one at a 1 m skin per body, though
    # their centres are two of the old grid's cells apart
    bodies = [dict(shape="box", size=[1, 1, 1], density=1, friction=0.5, pos=[0, 0, 0]),
              dict(shape="box", size=[1, 1, 1], density=1, friction=0.5, pos=[3.6, 0, 0])]
    g = avbd.HipBackend(precision=avbd.PREC_EXACT)
    g.pack(dict(bodies=bodies, joints=[]), dict(GPU, gravity=0.0))
    g.step(1)
    assert len(g.read_pairs()) == 0
    g.set_params(contactOffset=1.0)
    g.step(1)
    assert g.read_pairs().tolist() == [[1, 0]]
    assert g.counters()["manifolds"] == 0  # 2.6 m apart: outside the 2 m skin
