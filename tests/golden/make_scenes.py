"""Writes the `.scene` fixtures under tests/golden/scenes/.

The scenes are the reference parity harness's set (test/parity/gen_scenes.mjs:
71-261), restated here as data through this package's own writer. Run from the repository root:

    python tests/golden/make_scenes.py
"""
import math
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "avbd-sandbox_amd"))
from avbd import sceneio  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "scenes")


def body(size, density, friction, pos, vel=(0, 0, 0), quat=(0, 0, 0, 1), angvel=(0, 0, 0)):
    return dict(shape="box", size=list(size), density=density, friction=friction, pos=list(pos), vel=list(vel),
                quat=list(quat), angvel=list(angvel))


def joint(a, b, rA, rB, kLin=math.inf, kAng=0.0, fracture=math.inf):
    return dict(a=a, b=b, rA=list(rA), rB=list(rB), kLin=kLin, kAng=kAng, fracture=fracture)


def ground():
    return dict(bodies=[body([100, 100, 1], 0, 0.5, [0, 0, 0])], joints=[])


def scenes():
    S = {}
    s = ground()
    s["bodies"].append(body([1, 1, 1], 1, 0.5, [0, 0, 4]))
    S["ground"] = s

    S["spinner"] = dict(bodies=[body([2, 0.6, 1.4], 1, 0.5, [0, 0, 0], angvel=[1.3, 2.1, -0.7])], joints=[])

    s = ground()
    for i in range(10):
        s["bodies"].append(body([1, 1, 1], 1, 0.5, [0, 0, i * 1.5 + 1.0]))
    S["stack"] = s

    s = ground()
    top, size = 0.5, 1.0
    for _ in range(4):
        half = size * 0.5
        cz = top + half
        s["bodies"].append(body([size] * 3, 1, 0.5, [0, 0, cz]))
        top, size = cz + half, size * 2.0
    S["stack_ratio"] = s

    s = dict(bodies=[body([100, 100, 1], 0, 0.5, [0, 0, -0.5])], joints=[])
    n = 8
    for y in range(n):
        for x in range(n - y):
            s["bodies"].append(body([1, 0.5, 0.5], 1, 0.5, [x * 1.01 + y * 0.5 - n / 2, 0, y * 0.85 + 0.5]))
    S["pyramid"] = s

    def rope(name, links, heavy=False, angular=False):
        s = dict(bodies=[body([100, 100, 1], 0, 0.5, [0, 0, -20])], joints=[])
        prev = -1
        for i in range(links):
            last = heavy and i == links - 1
            big = 5
            s["bodies"].append(body([big] * 3 if last else [1, 0.5, 0.5], 0 if i == 0 else 1, 0.5,
                                    [i + (big / 2 if last else 0), 0, 10]))
            cur = len(s["bodies"]) - 1
            if prev >= 0:
                s["joints"].append(joint(prev, cur, [0.5, 0, 0], [-big / 2, 0, 0] if last else [-0.5, 0, 0],
                                         kAng=math.inf if angular else 0.0))
            prev = cur
        S[name] = s
    rope("rope", 20)
    rope("heavy_rope", 20, heavy=True)
    rope("rope_angular", 12, angular=True)

    s = ground()
    angle = 30 * math.pi / 180
    h = angle * 0.5
    q = [0, math.sin(h), 0, math.cos(h)]
    s["bodies"].append(body([40, 24, 1], 0, 1.0, [0, 0, 3], quat=q))

    def rot(v):
        x, y, z, w = q
        tx = 2 * (y * v[2] - z * v[1])
        ty = 2 * (z * v[0] - x * v[2])
        tz = 2 * (x * v[1] - y * v[0])
        return [v[0] + tx * w + (y * tz - z * ty), v[1] + ty * w + (z * tx - x * tz),
                v[2] + tz * w + (x * ty - y * tx)]
    t, nrm = rot([1, 0, 0]), rot([0, 0, 1])
    for i in range(11):
        mu = (i / 10) * 0.25 + 0.25
        y = -10 + i * 2
        s["bodies"].append(body([1, 1, 1], 1, mu, [0 + t[0] * -12 + 0 + nrm[0] * 1.05,
                                                   0 + t[1] * -12 + y + nrm[1] * 1.05,
                                                   3 + t[2] * -12 + 0 + nrm[2] * 1.05]))
    S["static_friction"] = s

    s = ground()
    for i in range(12):
        a = 0.3 + i * 0.37
        nn = math.hypot(0.3, 0.8, 0.5)
        sn = math.sin(a * 0.5)
        s["bodies"].append(body([1.4, 0.8, 0.6], 1, 0.4,
                                [(i % 3) * 1.3 - 1.3, math.floor(i / 3) * 1.3 - 1.3, 2 + i * 1.1],
                                vel=[0.4, -0.2, 0],
                                quat=[0.3 / nn * sn, 0.8 / nn * sn, 0.5 / nn * sn, math.cos(a * 0.5)],
                                angvel=[0.5, -0.9, 0.3]))
    S["tumble"] = s

    def spring(a, b, rA, rB, stiffness, rest):
        return dict(type="spring", a=a, b=b, rA=list(rA), rB=list(rB), stiffness=stiffness, rest=rest,
                    tearStrain=math.inf)
    s = ground()
    s["bodies"].append(body([1, 1, 1], 0, 0.5, [0, 0, 14]))
    s["bodies"].append(body([2, 2, 2], 1, 0.5, [0, 0, 8]))
    s["joints"].append(spring(1, 2, [0, 0, 0], [0, 0, 0], 100, 4))
    S["spring"] = s

    s = dict(bodies=[body([100, 100, 1], 0, 0.5, [0, 0, -10])], joints=[])
    prev = -1
    for i in range(8):
        s["bodies"].append(body([1, 0.75, 0.75], 0 if i in (0, 7) else 1, 0.5, [(i - 3.5) * 3.0, 0, 12]))
        cur = len(s["bodies"]) - 1
        if prev >= 0:
            s["joints"].append(spring(prev, cur, [0.5, 0, 0], [-0.5, 0, 0], 10 if i % 2 == 0 else 10000, 3))
        prev = cur
    S["spring_ratio"] = s

    # a spring of rest -1 takes the anchors' distance when it is made
    # (spring.js:79-84), beside one of a given rest; off-centre anchors
    s = ground()
    s["bodies"].append(body([1, 1, 1], 0, 0.5, [0, 0, 10]))
    s["bodies"].append(body([1, 0.5, 0.5], 1, 0.5, [0.7, 0.2, 7.3]))
    s["bodies"].append(body([0.6, 0.6, 0.6], 1, 0.5, [2.1, -0.4, 5.9]))
    s["joints"].append(spring(1, 2, [0.25, 0, -0.5], [0.5, 0.1, 0], 300, -1))
    s["joints"].append(spring(2, 3, [0, 0, -0.25], [0, 0, 0.3], 150, 1.5))
    S["spring_current"] = s
    return S


if __name__ == "__main__":
    os.makedirs(OUT, exist_ok=True)
    for name, sc in scenes().items():
        sceneio.write_scene(sc, os.path.join(OUT, name + ".scene"), name)
        print(name, len(sc["bodies"]), "bodies", len(sc["joints"]), "joints")
