"""The `.scene` interchange format and the %.17g trajectory dump.

A `.scene` file is the record list the reference's parity harness feeds to
both of its engines (format: test/parity/gen_scenes.mjs:8-14; readers:
test/parity/scene_loader.mjs:20-84 and test/parity/reference_main.cpp:45-124):

    body   sx sy sz density friction  px py pz  vx vy vz  qx qy qz qw  wx wy wz
    joint  idxA idxB  rax ray raz  rbx rby rbz  stiffLin stiffAng fracture
    spring idxA idxB  rax ray raz  rbx rby rbz  stiffness rest

'#' starts a comment; tokens are separated by blanks, tabs and CR; "inf" and
"-inf" are the infinities. Bodies are boxes, numbered in creation order; a
joint's idxA of -1 anchors it to the world (rA is then a world point).
Joints and springs are forces in file order; a spring's negative rest length
means the anchors' current distance.

The trajectory dump is reference_main.cpp's output (lines 126-140): after each
step, one line per body, "step index px py pz qx qy qz qw vx vy vz wx wy wz",
numbers as C's %.17g, so that every f64 survives the round trip exactly.
"""
import math

import numpy as np

from . import _lib as L
from .backend import AvbdError, HipBackend

BODY_FIELDS = 18
JOINT_FIELDS = 11  # idxA idxB + 9 numbers
SPRING_FIELDS = 10  # idxA idxB + 8 numbers


def _value(tok):
    """One number token (reference_main.cpp:36-43, scene_loader.mjs:14-18)."""
    if tok == "inf":
        return math.inf
    if tok == "-inf":
        return -math.inf
    try:
        return float(tok)
    except ValueError:
        raise AvbdError(L.AVBD_E_INVALID, f"not a number: {tok!r}") from None


def parse_scene(text):
    """Scene dict (the description HipBackend.pack takes) of a `.scene` text."""
    bodies, joints = [], []
    for lineno, raw in enumerate(text.split("\n"), 1):
        tok = raw.split("#", 1)[0].replace("\t", " ").replace("\r", " ").split()
        if not tok:
            continue
        kind = tok[0]
        if kind == "body":
            if len(tok) < 1 + BODY_FIELDS:
                raise AvbdError(L.AVBD_E_INVALID, f"line {lineno}: body record needs {BODY_FIELDS} numbers")
            v = [_value(t) for t in tok[1:1 + BODY_FIELDS]]
            bodies.append(dict(shape="box", size=v[0:3], density=v[3], friction=v[4], pos=v[5:8],
                               vel=v[8:11], quat=v[11:15], angvel=v[15:18]))
        elif kind == "joint":
            if len(tok) < 1 + JOINT_FIELDS:
                raise AvbdError(L.AVBD_E_INVALID, f"line {lineno}: joint record needs {JOINT_FIELDS} fields")
            # (int) of the parsed value, as the C reader does
            ia, ib = int(_value(tok[1])), int(_value(tok[2]))
            if not (0 <= ib < len(bodies)) or ia >= len(bodies):
                raise AvbdError(L.AVBD_E_INVALID, f"line {lineno}: joint names a body not yet defined")
            v = [_value(t) for t in tok[3:12]]
            joints.append(dict(a=ia if ia >= 0 else -1, b=ib, rA=v[0:3], rB=v[3:6], kLin=v[6], kAng=v[7],
                               fracture=v[8]))
        elif kind == "spring":
            if len(tok) < 1 + SPRING_FIELDS:
                raise AvbdError(L.AVBD_E_INVALID, f"line {lineno}: spring record needs {SPRING_FIELDS} fields")
            ia, ib = int(_value(tok[1])), int(_value(tok[2]))
            if not (0 <= ia < len(bodies)) or not (0 <= ib < len(bodies)):
                raise AvbdError(L.AVBD_E_INVALID, f"line {lineno}: spring names a body not yet defined")
            v = [_value(t) for t in tok[3:11]]
            joints.append(dict(type="spring", a=ia, b=ib, rA=v[0:3], rB=v[3:6], stiffness=v[6], rest=v[7],
                               tearStrain=math.inf))
        else:
            raise AvbdError(L.AVBD_E_INVALID, f"line {lineno}: unknown record {kind!r}")
    return dict(bodies=bodies, joints=joints)


def read_scene(path):
    with open(path, "r") as f:
        return parse_scene(f.read())


def format_number(v):
    """A number as the scene writer emits it (gen_scenes.mjs:57-64): the
    infinities as "inf"/"-inf", integers without a fraction, anything else
    with 17 significant digits, which reads back to the same f64. Negative
    zero is written "-0" so that it, too, reads back unchanged."""
    v = float(v)
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    if math.isnan(v):
        raise AvbdError(L.AVBD_E_INVALID, "NaN cannot be written to a scene")
    if v == 0.0:
        return "-0" if math.copysign(1.0, v) < 0 else "0"
    if v.is_integer() and abs(v) < 1e21:
        return str(int(v))
    return "%.17g" % v


def _num(x):
    if isinstance(x, str):
        return {"inf": math.inf, "-inf": -math.inf}[x]
    return float(x)


def format_scene(scene, name=None):
    """`.scene` text of a scene dict (boxes and joints; gen_scenes.mjs:25-55)."""
    lines = [f"# scene: {name}"] if name else []
    for i, b in enumerate(scene["bodies"]):
        if b.get("shape", "box") != "box":
            raise AvbdError(L.AVBD_E_INVALID, f"body {i}: the .scene format holds boxes only")
        if b.get("scene", 0) != 0:
            raise AvbdError(L.AVBD_E_INVALID, f"body {i}: the .scene format holds one scene")
        vals = (list(b["size"]) + [b["density"], b["friction"]] + list(b["pos"]) +
                list(b.get("vel", (0, 0, 0))) + list(b.get("quat", (0, 0, 0, 1))) +
                list(b.get("angvel", (0, 0, 0))))
        lines.append(" ".join(["body"] + [format_number(x) for x in vals]))
    for j in scene.get("joints", []):
        if j.get("type") == "spring":
            if not math.isinf(_num(j.get("tearStrain", "inf"))):
                raise AvbdError(L.AVBD_E_INVALID, "the .scene format has no tearing strain")
            vals = list(j["rA"]) + list(j["rB"]) + [_num(j["stiffness"]), _num(j.get("rest", -1.0))]
            lines.append(" ".join(["spring", str(int(j["a"])), str(int(j["b"]))] +
                                  [format_number(x) for x in vals]))
            continue
        vals = list(j["rA"]) + list(j["rB"]) + [_num(j["kLin"]), _num(j["kAng"]), _num(j["fracture"])]
        lines.append(" ".join(["joint", str(int(j["a"])), str(int(j["b"]))] + [format_number(x) for x in vals]))
    return "\n".join(lines) + "\n"


def write_scene(scene, path, name=None):
    with open(path, "w") as f:
        f.write(format_scene(scene, name))


def format_frame(step, state):
    """One frame of the trajectory dump (reference_main.cpp:126-140) from a
    state dict as HipBackend.read_state returns it."""
    cols = np.concatenate([np.asarray(state["pos"], np.float64), np.asarray(state["quat"], np.float64),
                           np.asarray(state["vel"], np.float64), np.asarray(state["angvel"], np.float64)],
                          axis=1)
    fmt = "%d %d " + " ".join(["%.17g"] * 13) + "\n"
    return "".join(fmt % ((step, i) + tuple(row)) for i, row in enumerate(cols.tolist()))


def parse_trajectory(text):
    """{step: (n, 13) array p q v w} of a dump (the reader of
    run_parity.mjs:126-139)."""
    frames = {}
    for line in text.splitlines():
        if not line:
            continue
        p = line.split(" ")
        step, idx = int(p[0]), int(p[1])
        frames.setdefault(step, {})[idx] = [float(x) for x in p[2:15]]
    return {s: np.array([rows[i] for i in range(len(rows))]) for s, rows in frames.items()}


def run_scene(path, steps=120, iterations=None, dt=None, gravity=None, alpha=None, beta_lin=None,
              beta_ang=None, gamma=None, precision=L.PREC_EXACT, out=None, backend=None):
    """Load a `.scene`, step it `steps` times on the GPU and write the
    trajectory dump after every step (reference_main.cpp:142-183; the
    optional parameters are its positional arguments, in its order; those
    not given keep Solver.defaultParams, solver.js:92-236). Returns the dump
    text, or writes it to the file object `out` and returns None."""
    scene = read_scene(path)
    params = {k: v for k, v in dict(iterations=iterations, dt=dt, gravity=gravity, alpha=alpha,
                                     betaLin=beta_lin, betaAng=beta_ang, gamma=gamma).items() if v is not None}
    gpu = backend if backend is not None else HipBackend(precision=precision)
    gpu.pack(scene, params)
    chunks = []
    for s in range(1, steps + 1):
        gpu.step(1)
        frame = format_frame(s, gpu.read_state())
        if out is not None:
            out.write(frame)
        else:
            chunks.append(frame)
    return None if out is not None else "".join(chunks)
