"""The pose export that allocates nothing (avbd_export_poses_static) against
avbd_export_poses, through the C ABI.

avbd_export_poses stages host-bound poses in a device buffer the context
allocates on first use; avbd_export_poses_static stages them in a static
device array instead, for callers that cannot allocate. The two must write the
same bits: on two calls in a row, the second on a smaller scene (the static
array keeps the first call's contents beyond the second's bodies), from two
contexts that share the device's array, and at both blend fractions. A scene
larger than the array is rejected before any launch.
"""
import ctypes as C

import numpy as np
import pytest

import avbd
from avbd import _lib as L

pytestmark = pytest.mark.gpu


def via(gpu, name, alpha):
    out = np.zeros((gpu.n, 8), np.float32)
    fn = getattr(gpu.L, name)
    args = (gpu.ctx, float(alpha), out.ctypes.data_as(C.c_void_p), gpu.n) + ((0,) if name == "avbd_export_poses" else ())
    rc = fn(*args)
    assert rc == 0, gpu.L.avbd_last_error(gpu.ctx).decode()
    return out


def test_static_export_equals_export_on_two_calls():
    ctxs = []
    for layers in (8, 5):  # the second call has fewer bodies than the first
        scene, params, _ = avbd.scenes.pyramid(layers=layers)
        gpu = avbd.HipBackend(precision=avbd.PREC_MIXED)
        gpu.pack(scene, params)
        gpu.step(3)
        ctxs.append(gpu)
        for alpha in (0.37, 1.0):
            a = via(gpu, "avbd_export_poses_static", alpha)
            b = via(gpu, "avbd_export_poses", alpha)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (layers, alpha)
    # both contexts share the device's array: calls in turn still agree
    for gpu in ctxs + ctxs[::-1]:
        gpu.step(1)
        a = via(gpu, "avbd_export_poses_static", 0.5)
        b = via(gpu, "avbd_export_poses", 0.5)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_static_export_rejects_a_scene_beyond_the_array():
    n = L.STATIC_POSE_BODIES + 1
    side = int(np.ceil(n ** 0.5))
    bodies = [avbd.scenes.sphere(0.5, 1.0, 0.5, [3.0 * (i % side), 3.0 * (i // side), 5.0]) for i in range(n)]
    gpu = avbd.HipBackend(precision=avbd.PREC_MIXED, manifold_cap=4096, pair_cap=4096)
    gpu.pack(dict(bodies=bodies, joints=[]), dict(gravity=0.0))
    out = np.zeros((n, 8), np.float32)
    rc = gpu.L.avbd_export_poses_static(gpu.ctx, 1.0, out.ctypes.data_as(C.c_void_p), n)
    assert rc == L.AVBD_E_INVALID and "at most" in gpu.L.avbd_last_error(gpu.ctx).decode()
    assert not out.any()
    # the context is unharmed: the allocating export still serves it
    assert np.array_equal(gpu.export_poses(1.0)[:, :3], gpu.read_state()["pos"].astype(np.float32))
