"""Poses for drawing: avbd_export_poses against the oracle's restatement of
the renderer's blend (renderer.js:1327-1383), bit for bit in f32, to host
memory and to device memory; avbd_pose_source's arena addresses read back
the current and step-start poses."""
import ctypes
import json
import os
import shutil
import subprocess

import numpy as np
import pytest
import torch

import avbd
import pyoracle
from avbd import AvbdError, sceneio

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def stepped(name="tumble", steps=25):
    g = avbd.HipBackend(precision=avbd.PREC_MIXED)
    g.pack(sceneio.read_scene(os.path.join(ROOT, "tests", "golden", "scenes", name + ".scene")))
    g.step(steps - 1)
    before = g.read_state()
    g.step(1)
    return g, before, g.read_state()


@pytest.mark.parametrize("alpha", [1.0, 0.0, 0.25, 0.5, 0.999, 1.5, -0.3, float("nan")])
def test_export_matches_the_oracle_blend(alpha):
    g, s0, s1 = stepped()
    want = pyoracle.render_poses(s1["pos"], s1["quat"], s0["pos"], s0["quat"], alpha)
    host = g.export_poses(alpha)
    assert np.array_equal(host, want)
    dev = torch.full((g.n, 8), float("nan"), dtype=torch.float32, device="cuda:0")
    # the fill runs on torch's stream and the export on the library's: the
    # destination must be ready before the export is ordered after it
    torch.cuda.synchronize()
    g.export_poses(alpha, out=dev)
    g.idle()
    assert np.array_equal(dev.cpu().numpy(), want)


def test_large_scene_blend_and_rejections():
    scene, params, _ = avbd.scenes.rain(n=20000)
    g = avbd.HipBackend(precision=avbd.PREC_MIXED)
    g.pack(scene, params)
    g.step(30)
    s0 = g.read_state()
    g.step(1)
    s1 = g.read_state()
    want = pyoracle.render_poses(s1["pos"], s1["quat"], s0["pos"], s0["quat"], 0.3)
    assert np.array_equal(g.export_poses(0.3), want)
    with pytest.raises(AvbdError):
        g.export_poses(0.3, out=np.zeros((g.n - 1, 8), np.float32))
    # a host address passed as device memory is refused, not written
    rc = g.L.avbd_export_poses(g.ctx, 0.5, np.zeros((g.n, 8), np.float32).ctypes.data_as(ctypes.c_void_p), g.n, 1)
    assert rc == avbd._lib.AVBD_E_INVALID
    assert g.L.avbd_export_poses(g.ctx, 0.5, None, g.n, 0) == avbd._lib.AVBD_E_INVALID


def test_pose_source_reads_current_and_step_start_pose():
    g, s0, s1 = stepped("rope_angular", 12)
    pose, start, stride = g.pose_source()
    assert stride == 64 and pose and start
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    g.idle()
    for addr, s in ((pose, s1), (start, s0)):
        buf = np.zeros((g.n, 8), np.float64)
        assert hip.hipMemcpy(buf.ctypes.data, addr, buf.nbytes, 2) == 0  # device to host
        assert np.array_equal(buf[:, 0:3], s["pos"]) and np.array_equal(buf[:, 3:7], s["quat"])


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_js_host_export_poses():
    path = os.path.join(ROOT, "tests", "golden", "scenes", "tumble.scene")
    r = subprocess.run([shutil.which("node"), "tests/js/run_scene.mjs", "poses", path, "20", "mixed"],
                       capture_output=True, text=True, cwd=ROOT, timeout=120)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    s0, s1 = np.array(d["s0"]), np.array(d["s1"])
    want = pyoracle.render_poses(s1[:, :3], s1[:, 3:], s0[:, :3], s0[:, 3:], 0.5)
    assert np.array_equal(np.array(d["poses"], np.float32).reshape(-1, 8), want)
