"""CPU checks of the C ABI library: it loads and exports every symbol that
include/avbd.h declares (no compute call is made without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "avbd.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(avbd_[a-z_0-9]+)\s*\(", text)))


def test_header_symbols_exported():
    import avbd
    if not os.path.exists(avbd.LIB_PATH):
        pytest.fail("libavbd_hip.so not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(avbd.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"not exported: {missing}"
    assert set(declared_symbols()) == set(avbd.SYMBOLS)


def test_abi_version_and_defaults():
    import avbd
    from avbd import _lib as L
    lib = avbd.lib()
    assert lib.avbd_abi_version() == 6
    p = L.Params()
    lib.avbd_default_params(ctypes.byref(p))
    # solver.js:92-236
    assert p.dt == 1.0 / 60.0 and p.gravity == -10.0 and p.iterations == 10
    assert p.beta_lin == 1e4 and p.beta_ang == 100.0 and p.alpha == 0.95 and p.gamma == 0.999
    assert p.penalty_min == 1.0 and p.penalty_max == 1e10 and p.collision_margin == 0.01
    assert p.stick_threshold == 1e-5 and p.post_stabilize == 0 and p.rotated_inertia == 1
    # the GPU contact law's parameters (solver.js:151-162), the CPU law selected
    assert p.contact_model == L.CONTACTS_CPU and p.contact_offset == 0.005 and p.rest_offset == -0.005
    assert p.restitution_threshold == 1.0 and p.contact_persistence == 0.002


def test_null_and_invalid_arguments():
    import avbd
    lib = avbd.lib()
    assert lib.avbd_destroy(None) == -1
    assert lib.avbd_step(None, 1) == -1
    assert lib.avbd_create(None, None) == -1


def test_scene_arrays_layout():
    import numpy as np
    from avbd import scenes, scene_arrays
    scene, params, meta = scenes.batched(num_scenes=6)
    A = scene_arrays(scene)
    n = len(scene["bodies"])
    assert A["pos"].shape == (n, 3) and A["quat"].shape == (n, 4)
    assert (np.diff(A["scene_of"]) >= 0).all()
    assert A["joint_a"].shape[0] == len(scene["joints"])
    # every joint stays inside its scene
    so = A["scene_of"]
    for a, b in zip(A["joint_a"], A["joint_b"]):
        assert a < 0 or so[a] == so[b]
