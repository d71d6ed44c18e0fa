"""Seeded synthetic scenes for the five benchmark configurations.

Each generator returns ``(scene, params, meta)``: the scene dict HipBackend.pack
and the oracle both take, solver parameters (reference names), and metadata
(steps, the body to launch and when, a description). Geometry follows the
reference's scene builders where the configuration leaves a choice open
(src/app/scenes.js: ground 97-103, box_stack 290-308, pendulum 391-439,
friction_ramp 519-562, great_pyramid 771-805); the seeded generator is the
LCG of scenes.js:78-84.
"""
import math


def lcg(seed):
    """Uniform [0, 1) stream: s <- (1664525 s + 1013904223) mod 2^32."""
    s = seed & 0xFFFFFFFF

    def r():
        nonlocal s
        s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
        return s / 4294967296.0
    return r


def box(size, density, friction, pos, **kw):
    d = dict(shape="box", size=list(size), density=density, friction=friction, pos=list(pos))
    d.update(kw)
    return d


def sphere(diameter, density, friction, pos, **kw):
    d = dict(shape="sphere", size=[diameter] * 3, density=density, friction=friction, pos=list(pos))
    d.update(kw)
    return d


def ground(friction=0.5, **kw):
    return box([400, 400, 4], 0, friction, [0, 0, -2], **kw)


def axis_angle(ax, ay, az, angle):
    n = math.sqrt(ax * ax + ay * ay + az * az)
    h = angle * 0.5
    s = math.sin(h)
    return [ax / n * s, ay / n * s, az / n * s, math.cos(h)]


def stack10(seed=1):
    """Config 1: 10 unit cubes (rho 1000, mu 0.5) at 1.001 m pitch on a static
    ground, each jittered horizontally by U(-0.01, 0.01); g = -9.81, 10
    iterations, 300 steps."""
    r = lcg(seed)
    bodies = [ground(0.5)]
    for i in range(10):
        jx = (r() * 2 - 1) * 0.01
        jy = (r() * 2 - 1) * 0.01
        bodies.append(box([1, 1, 1], 1000, 0.5, [jx, jy, 0.501 + i * 1.001]))
    return dict(bodies=bodies, joints=[]), dict(gravity=-9.81, iterations=10), dict(steps=300, seed=seed)


def pyramid(layers=53, density=1.0, friction=0.6, sphere_density=10.0, speed=20.0):
    """Config 2: square pyramid of unit cubes at 1 m pitch, layer k (from the
    top) holding k x k cubes: 53 layers = 51,039 cubes. A 1 m-diameter sphere
    of 10x the cube density is parked on the ground far away and, after 60
    settling steps, launched at the apex at 20 m/s (``meta['launch']``)."""
    bodies = [ground(friction)]
    for layer in range(layers):
        w = layers - layer
        off = (w - 1) / 2.0
        for x in range(w):
            for y in range(w):
                bodies.append(box([1, 1, 1], density, friction, [x - off, y - off, layer + 0.5]))
    ball = len(bodies)
    bodies.append(sphere(1.0, density * sphere_density, friction, [150.0, 150.0, 0.5]))
    launch = dict(body=ball, step=60, pos=[0.0, 0.0, layers + 2.0], vel=[0.0, 0.0, -speed])
    return (dict(bodies=bodies, joints=[]), dict(gravity=-9.81, iterations=20),
            dict(steps=600, launch=launch, cubes=ball - 1))


def rain(n=50000, seed=7, width=40.0, height=100.0):
    """Config 3: n bodies, 70% boxes (edges U[0.2, 1.0] per axis), 30% spheres
    (radius U[0.1, 0.5]), density log-uniform over [1, 1000], friction
    U[0.2, 0.9], random orientation, positions uniform in a width x width x
    height column above a floor and four static walls; 20 iterations."""
    r = lcg(seed)
    h = width / 2.0
    bodies = [ground(0.6)]
    wall_h = 8.0
    bodies.append(box([1.0, width + 2, wall_h], 0, 0.5, [-h - 0.5, 0, wall_h / 2]))
    bodies.append(box([1.0, width + 2, wall_h], 0, 0.5, [h + 0.5, 0, wall_h / 2]))
    bodies.append(box([width + 2, 1.0, wall_h], 0, 0.5, [0, -h - 0.5, wall_h / 2]))
    bodies.append(box([width + 2, 1.0, wall_h], 0, 0.5, [0, h + 0.5, wall_h / 2]))
    for _ in range(n):
        pos = [(r() * 2 - 1) * (h - 0.6), (r() * 2 - 1) * (h - 0.6), 1.0 + r() * height]
        density = math.exp(r() * math.log(1000.0))
        friction = 0.2 + r() * 0.7
        q = axis_angle(r() - 0.5, r() - 0.5, r() - 0.5 + 1e-3, r() * 6.283185307179586)
        if r() < 0.7:
            bodies.append(box([0.2 + r() * 0.8, 0.2 + r() * 0.8, 0.2 + r() * 0.8], density, friction, pos, quat=q))
        else:
            bodies.append(sphere(2.0 * (0.1 + r() * 0.4), density, friction, pos, quat=q))
    return dict(bodies=bodies, joints=[]), dict(gravity=-9.81, iterations=20), dict(steps=1000, seed=seed)


def chain(links, rng_angle=40.0, link=0.5, anchor_z=None):
    """A hanging chain of hard ball-socket links (pendulum, scenes.js:391-439):
    static anchor link, links-1 dynamic links (rho 1), and a 3 m cube of
    rho 60 at the end (mass ratio 25,920:1 to a link)."""
    if anchor_z is None:
        anchor_z = links * link + 6.0
    angle = math.radians(rng_angle)
    q = axis_angle(0, 1, 0, angle)

    def along(s):
        return [-math.sin(angle) * s, 0.0, anchor_z - math.cos(angle) * s]
    bodies, joints = [ground(0.5)], []
    for i in range(links):
        bodies.append(box([link, link * 0.5, link], 0 if i == 0 else 1, 0.4, along(i * link), quat=q))
        if i > 0:
            joints.append(dict(a=i, b=i + 1, rA=[0, 0, -link / 2], rB=[0, 0, link / 2], kLin="inf", kAng=0,
                               fracture="inf"))
    bodies.append(box([3, 3, 3], 60, 0.5, along((links - 1) * link + link / 2 + 1.5), quat=q))
    joints.append(dict(a=links, b=links + 1, rA=[0, 0, -link / 2], rB=[0, 0, 1.5], kLin="inf", kAng=0,
                       fracture="inf"))
    return bodies, joints


def stack(cubes, seed):
    r = lcg(seed)
    bodies = [ground(0.5)]
    for i in range(cubes):
        bodies.append(box([1, 1, 1], 1, 0.5, [(r() * 2 - 1) * 0.01, (r() * 2 - 1) * 0.01, 0.501 + i * 1.001]))
    return bodies, []


def batched(num_scenes=256, seed=3):
    """Config 4: num_scenes independent scenes; even ones are chain pendulums of
    U{10..200} hard ball-socket links with a 26,000:1 end mass, odd ones box
    stacks of U{5..50} cubes; 20 iterations, 300 steps. Bodies carry their
    scene id; scenes never interact. Returns the merged scene and, in meta,
    the per-scene scene dicts (for running the oracle scene by scene)."""
    r = lcg(seed)
    bodies, joints, parts = [], [], []
    for s in range(num_scenes):
        if s % 2 == 0:
            b, j = chain(10 + int(r() * 191))
        else:
            b, j = stack(5 + int(r() * 46), seed * 1000 + s)
        parts.append(dict(bodies=b, joints=j))
        base = len(bodies)
        for x in b:
            y = dict(x)
            y["scene"] = s
            bodies.append(y)
        for x in j:
            y = dict(x)
            y["a"] = x["a"] + base if x["a"] >= 0 else -1
            y["b"] = x["b"] + base
            joints.append(y)
    return (dict(bodies=bodies, joints=joints), dict(gravity=-9.81, iterations=20),
            dict(steps=300, parts=parts, seed=seed))


def ramp(n=64, seed=5, angle_deg=30.0):
    """Config 5: n boxes (0.5 m, rho 500) on a static 30-degree ramp, friction
    U[0.1, 0.9], initial speed U[0, 2] m/s down-slope; 10 iterations, 120
    steps. meta['targets'] holds the loss's seeded targets, one per body
    (the statics' are their own positions): each box's start moved U[0, 3] m
    down the slope and U[-0.5, 0.5] m across it; meta['slope'] is the unit
    down-slope direction."""
    r = lcg(seed)
    a = math.radians(angle_deg)
    q = axis_angle(0, 1, 0, a)
    bodies = [ground(0.5), box([40, 24, 1], 0, 1.0, [0, 0, 10], quat=q)]
    c, s = math.cos(a), math.sin(a)
    per_row = 8
    for i in range(n):
        u = -14 + (i % per_row) * 2.0
        w = -9 + (i // per_row) * 2.5
        hgt = 0.5 + 0.25 + 0.001
        pos = [c * u + s * hgt, w, 10 - s * u + c * hgt]
        v0 = r() * 2
        bodies.append(box([0.5, 0.5, 0.5], 500, 0.1 + r() * 0.8, pos, quat=q, vel=[c * v0, 0, -s * v0]))
    t = lcg(seed * 7919 + 1)
    targets = [list(b["pos"]) for b in bodies[:2]]
    for b in bodies[2:]:
        d, w = 3.0 * t(), t() - 0.5
        targets.append([b["pos"][0] + c * d, b["pos"][1] + w, b["pos"][2] - s * d])
    return (dict(bodies=bodies, joints=[]), dict(gravity=-9.81, iterations=10),
            dict(steps=120, seed=seed, targets=targets, slope=[c, 0.0, -s]))


def ramp_lanes(n=8, seed=3, angle_deg=30.0):
    """n boxes (0.5 m, rho 500) side by side across the 30-degree ramp, one per
    lane 2.5 m apart so that no two touch, friction U[0.1, 0.4] (all below
    tan 30 deg: they keep sliding) and initial speed U[0, 2] m/s down-slope.
    The smooth sliding regime in which finite differences check gradients."""
    r = lcg(seed)
    a = math.radians(angle_deg)
    q = axis_angle(0, 1, 0, a)
    c, s = math.cos(a), math.sin(a)
    bodies = [ground(0.5), box([40, 24, 1], 0, 1.0, [0, 0, 10], quat=q)]
    for i in range(n):
        u, w = -8.0, -9 + i * 2.5
        hgt = 0.5 + 0.25 + 0.001
        pos = [c * u + s * hgt, w, 10 - s * u + c * hgt]
        mu = 0.1 + 0.3 * r()
        v0 = 2 * r()
        bodies.append(box([0.5, 0.5, 0.5], 500, mu, pos, quat=q, vel=[c * v0, 0, -s * v0]))
    return dict(bodies=bodies, joints=[]), dict(gravity=-9.81, iterations=10), dict(slope=[c, 0.0, -s])


CONFIGS = {"stack10": stack10, "pyramid": pyramid, "rain": rain, "batched": batched, "ramp": ramp}


def fabric(count=15000, cell=0.9, density=1.0, friction=0.5, tear_strain=0.75, pins=2, balls=16, seed=9):
    """A hanging spring-lattice cloth of `count` sphere particles, the
    sandbox's fabric spawner (scenes.js:1185-1301): cols = ceil(sqrt(1.5
    count)), 0.55-cell beads, structural links (k = 240 density (cell/0.9)^2,
    scenes.js:989-990), shear links (0.7 k) and bend links (0.22 k), all of
    their initial length and tearing at strain `tear_strain`; `pins` top-edge
    particles static. `balls` heavy spheres (3 cells, 20x the density) are
    thrown at the curtain at 8 m/s from 5 cells in front of it."""
    cols = min(count, max(1, math.ceil(math.sqrt(count * 1.5))))
    rows = max(1, math.ceil(count / cols))
    pinned = set()
    for p in range(min(pins, cols)):
        pinned.add(0 if pins == 1 else round(p * (cols - 1) / (pins - 1)))
    diameter = cell * 0.55
    k = 240.0 * max(0.05, density) * (cell / 0.9) ** 2
    top = diameter + (rows - 1) * cell + 1.0
    bodies, joints = [ground(friction)], []
    grid = {}
    made = 0
    for r in range(rows):
        for c in range(cols):
            if made == count:
                break
            u = c / max(1, cols - 1)
            v = r / max(1, rows - 1)
            billow = math.sin(math.pi * u) * math.sin(math.pi * v) * cell * 0.18
            pin = r == 0 and c in pinned
            bodies.append(sphere(diameter, 0 if pin else density, friction,
                                 [(c - (cols - 1) * 0.5) * cell, billow, top - r * cell]))
            me = len(bodies) - 1
            grid[(r, c)] = me
            made += 1
            for (dr, dc, scale) in ((0, -1, 1.0), (-1, 0, 1.0), (-1, -1, 0.7), (-1, 1, 0.7), (0, -2, 0.22),
                                    (-2, 0, 0.22)):
                other = grid.get((r + dr, c + dc))
                if other is not None:
                    joints.append(dict(type="spring", a=other, b=me, rA=[0, 0, 0], rB=[0, 0, 0],
                                       stiffness=k * scale, rest=-1.0, tearStrain=tear_strain))
    r = lcg(seed)
    width = (cols - 1) * cell
    height = (rows - 1) * cell
    for i in range(balls):
        bodies.append(sphere(3 * cell, 20 * density, friction,
                             [(r() - 0.5) * width * 0.8, -5.0 * cell - i * 4 * cell, top - (0.2 + 0.6 * r()) * height],
                             vel=[0.0, 8.0, 0.0]))
    return (dict(bodies=bodies, joints=joints), dict(gravity=-9.81, iterations=10),
            dict(steps=300, rows=rows, cols=cols, springs=len(joints)))

