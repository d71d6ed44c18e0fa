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
from contact_model_cases import (box, check_incline, check_restitution_target, check_rest_offset_equilibrium,
                                 check_sliding_deceleration, check_speculative_skin, check_staged_restitution,
                                 check_static_and_dynamic_friction)


class OracleWorld:
    """The oracle behind the checks' world interface."""

    def __init__(self, bodies, params):
        self.o = pyoracle.Oracle(dict(bodies=bodies, joints=[]), params,
                                 color_order=int(params.get("colorOrder", 0)))

    def step(self, n=1):
        self.o.step(n)

    def state(self):
        return self.o.state()

    def manifolds(self):
        kab, _ = self.o.forces()
        return [(int(a), int(b), self.o.manifold(i)) for i, (k, a, b, _) in enumerate(kab) if k == 0]

    def set_params(self, **kw):
        self.o.set_params(**kw)

    def set_body(self, i, vel):
        self.o.set_body(i, vel=vel)


def world(bodies, **params):
    return pyoracle.Oracle(dict(bodies=bodies, joints=[]), dict(params, contactModel=1))


def test_restitution_is_a_one_step_impact_target():
    check_restitution_target(OracleWorld)


def test_restitution_stays_armed_across_a_speculative_only_step():
    check_staged_restitution(OracleWorld)


def test_speculative_contacts_and_the_skin():
    check_speculative_skin(OracleWorld)


def test_rest_offset_sets_the_equilibrium_independently_of_the_skin():
    check_rest_offset_equilibrium(OracleWorld)


def test_static_and_dynamic_friction():
    check_static_and_dynamic_friction(OracleWorld)


@pytest.mark.parametrize("mu_d", [0.2, 0.4])
def test_sliding_box_decelerates_at_mu_d_g(mu_d):
    check_sliding_deceleration(OracleWorld, mu_d)


def test_incline_static_friction_holds_and_dynamic_friction_slides():
    check_incline(OracleWorld)


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
