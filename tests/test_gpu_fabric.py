"""The sandbox's spring-lattice fabric (scenes.js:1185-1301) with spheres
thrown through it: in exact precision the GPU tracks the oracle bit for bit
while links tear; in mixed precision it tracks it within the pose tolerance
until the first contact-feature divergence."""
import numpy as np
import pytest

import avbd
import pyoracle

pytestmark = pytest.mark.gpu


def test_fabric_tearing_exact():
    scene, params, meta = avbd.scenes.fabric(count=150, balls=3, tear_strain=0.08)
    print(f"[fabric] {meta['rows']} x {meta['cols']}, {meta['springs']} springs", flush=True)
    gpu = avbd.HipBackend(precision=avbd.PREC_EXACT)
    gpu.pack(scene, params)
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR)
    for s in range(1, 61):
        gpu.step(1)
        ref.step(1)
        st = gpu.read_state()
        g = np.concatenate([st["pos"], st["quat"], st["vel"], st["angvel"]], axis=1)
        o = ref.state()[:, :13]
        assert np.array_equal(g, o), f"step {s}: max |diff| {np.abs(g - o).max():.3e}"
    gk, _ = gpu.read_forces()
    rk, _ = ref.forces()
    assert np.array_equal(gk, rk)
    js = gpu.read_joints()
    torn = int(js[:, 12].sum())
    print(f"[fabric] {torn} links torn after 60 steps", flush=True)
    assert torn > 0
    springs = [ref.spring(i)[2] for i, (k, *_r) in enumerate(rk) if k == 2]
    assert np.array_equal(js[:, 12], np.array(springs))


def test_fabric_mixed_tracks_oracle():
    scene, params, _ = avbd.scenes.fabric(count=300, balls=2)
    gpu = avbd.HipBackend(precision=avbd.PREC_MIXED)
    gpu.pack(scene, params)
    ref = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR)
    worst = 0.0
    for s in range(1, 41):
        gpu.step(1)
        ref.step(1)
        gk, _ = gpu.read_forces()
        rk, _ = ref.forces()
        if not np.array_equal(gk, rk):
            break
        err = np.abs(gpu.read_state()["pos"] - ref.state()[:, :3]).max()
        worst = max(worst, err)
        assert err <= 5e-4, f"step {s}: pose error {err:.3e}"
    print(f"[fabric] mixed: worst pose error {worst:.2e} over {s} steps", flush=True)
