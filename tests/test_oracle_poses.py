"""The oracle's restatement of the renderer's pose blend (renderer.js:
1327-1383), pinned by what the mathematics fixes: the end points, the
closed form of nlerp about one axis, the shorter-arc sign rule, linearity of
the position blend and the clamping of alpha."""
import math

import numpy as np

import pyoracle


def zrot(theta):
    return [0.0, 0.0, math.sin(theta / 2), math.cos(theta / 2)]


def test_end_points():
    rng = np.random.default_rng(3)
    p, p0 = rng.normal(size=(50, 3)) * 10, rng.normal(size=(50, 3)) * 10
    q, q0 = rng.normal(size=(50, 4)), rng.normal(size=(50, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    q0 /= np.linalg.norm(q0, axis=1)[:, None]
    one = pyoracle.render_poses(p, q, p0, q0, 1.0)
    assert np.array_equal(one[:, :3], p.astype(np.float32)) and np.array_equal(one[:, 4:], q.astype(np.float32))
    assert not one[:, 3].any()
    zero = pyoracle.render_poses(p, q, p0, q0, 0.0)
    assert np.array_equal(zero[:, :3], p0.astype(np.float32))
    # at a = 0 the orientation is q0 up to sign (the sign of the shorter arc)
    sgn = np.sign((q0 * q).sum(axis=1))[:, None]
    assert np.abs(zero[:, 4:] - (q0 * sgn)).max() < 1e-7
    # alpha clamps to [0, 1]; NaN draws the current pose
    assert np.array_equal(pyoracle.render_poses(p, q, p0, q0, 7.0), one)
    assert np.array_equal(pyoracle.render_poses(p, q, p0, q0, -2.0), zero)
    assert np.array_equal(pyoracle.render_poses(p, q, p0, q0, float("nan")), one)


def test_nlerp_closed_form_about_one_axis():
    # q0 = identity, q = rotation by theta about z: nlerp is the rotation by
    # phi about z with tan(phi / 2) = a sin(theta/2) / (1 - a + a cos(theta/2))
    for theta in (0.1, 1.0, 2.5, 3.0):
        for a in (0.1, 0.25, 0.5, 0.9):
            got = pyoracle.render_poses([[0, 0, 0]], [zrot(theta)], [[0, 0, 0]], [[0, 0, 0, 1]], a)[0, 4:]
            phi = 2 * math.atan2(a * math.sin(theta / 2), 1 - a + a * math.cos(theta / 2))
            assert np.abs(got - np.array(zrot(phi), np.float32)).max() <= 6e-8, (theta, a)


def test_shorter_arc_and_position_blend():
    q0 = [[0.0, 0.0, 0.0, 1.0]]
    q = zrot(1.2)
    a = pyoracle.render_poses([[2, 4, 6]], [q], [[0, 0, 0]], q0, 0.5)
    b = pyoracle.render_poses([[2, 4, 6]], [[-x for x in q]], [[0, 0, 0]], q0, 0.5)
    # -q is the same orientation as q: the blend takes the same (shorter)
    # arc, so the result is the same rotation, written with the other sign;
    # without the sign rule it would pass through the far side
    assert np.array_equal(b[0, 4:], -a[0, 4:])
    assert np.abs(a[0, 4:] - np.array(zrot(0.6), np.float32)).max() < 1e-7
    assert np.array_equal(a[0, :3], np.float32([1, 2, 3]))
