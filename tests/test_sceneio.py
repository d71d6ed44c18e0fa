"""The `.scene` format and the %.17g trajectory dump (CPU).

The package's reader is checked against the oracle's C restatement of the
reference loader (test/parity/reference_main.cpp:45-124): both read every
fixture under tests/golden/scenes/ and the two worlds must step to the same
bits. The dump writer is checked against C's printf %.17g (the oracle's
restatement of reference_main.cpp:126-140), including the numbers where
%.17g formatting has corner cases.
"""
import glob
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import avbd
import pyoracle
from avbd import AvbdError, sceneio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCENES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "scenes", "*.scene")))
NAMES = [os.path.splitext(os.path.basename(p))[0] for p in SCENES]


def test_fixture_set_is_the_reference_parity_set():
    # test/parity/gen_scenes.mjs:71-261, and spring_current: a spring of rest
    # -1 (the anchors' distance when it is made, spring.js:79-84)
    assert NAMES == sorted(["ground", "spinner", "stack", "stack_ratio", "pyramid", "rope", "heavy_rope",
                            "rope_angular", "static_friction", "spring", "spring_ratio", "tumble",
                            "spring_current"])


@pytest.mark.parametrize("path", SCENES, ids=NAMES)
def test_reader_matches_the_c_loader(path):
    scene = sceneio.read_scene(path)
    a = pyoracle.Oracle(scene, schedule=pyoracle.SCHED_COLOR)
    b = pyoracle.Oracle(path, schedule=pyoracle.SCHED_COLOR)
    assert a.n == b.n == len(scene["bodies"])
    assert np.array_equal(a.state(), b.state())
    assert np.array_equal(a.mass(), b.mass())
    a.step(30)
    b.step(30)
    assert np.array_equal(a.state(), b.state())
    assert a.format_state(30) == b.format_state(30)


@pytest.mark.parametrize("path", SCENES, ids=NAMES)
def test_write_read_round_trip(path):
    text = open(path).read()
    scene = sceneio.parse_scene(text)
    assert sceneio.format_scene(scene, NAMES[SCENES.index(path)]) == text
    again = sceneio.parse_scene(sceneio.format_scene(scene))
    assert again == scene


def test_numbers_round_trip_exactly():
    vals = [0.0, -0.0, 1.0, -3.0, 0.1, 1.0 / 3.0, 2.0 ** -1074, 2.2250738585072014e-308, 1e-7, 123456.789,
            1e21, 1e22, -1e300, 9007199254740993.0, math.inf, -math.inf, 0.25881904510252074]
    scene = dict(bodies=[dict(shape="box", size=vals[:3], density=vals[3], friction=vals[4], pos=vals[5:8],
                              vel=vals[8:11], quat=vals[11:15], angvel=vals[15:17] + [0.5])], joints=[])
    back = sceneio.parse_scene(sceneio.format_scene(scene))["bodies"][0]
    flat = back["size"] + [back["density"], back["friction"]] + back["pos"] + back["vel"] + back["quat"] + \
        back["angvel"][:2]
    for x, y in zip(vals, flat):
        assert x == y and math.copysign(1, x) == math.copysign(1, y), (x, y)


def test_comments_blanks_tabs_and_crlf():
    text = ("# a scene\r\n\r\n   \n"
            "body\t1 2 3  0 0.5  0 0 -1  0 0 0  0 0 0 1  0 0 0 # ground\r\n"
            "body 1 1 1 1 0.5 0 0 4 0.5 0 0 0 0 0 1 0 0 0 7 8 9\n"   # extra tokens are ignored
            "joint -1 1  0 0 5  0 0 0.5  inf 0 -inf\n")
    s = sceneio.parse_scene(text)
    assert [b["size"] for b in s["bodies"]] == [[1, 2, 3], [1, 1, 1]]
    assert s["bodies"][1]["vel"] == [0.5, 0, 0]
    j = s["joints"][0]
    assert j["a"] == -1 and j["b"] == 1 and j["rA"] == [0, 0, 5] and j["kLin"] == math.inf
    assert j["fracture"] == -math.inf


@pytest.mark.parametrize("text,c_code", [
    ("body 1 1 1 1 0.5 0 0 0 0 0 0 0 0 0 1 0 0 0\nspring 0 1 0 0 0 0 0 0 100 4\n", -2),
    ("body 1 1 1 1 0.5 0 0 0 0 0 0 0 0 0 1 0 0 0\nspring 0 0 0 0 0 0 0 100\n", -2),
    ("box 1 1 1\n", -4),
    ("body 1 1 1 1 0.5 0 0 0\n", -2),
    ("body 1 1 1 1 0.5 0 0 0 0 0 0 0 0 0 1 0 0 0\njoint 0 1 0 0 0 0 0 0 inf 0 inf\n", -2),
], ids=["spring-forward-reference", "short-spring", "unknown", "short-body", "joint-forward-reference"])
def test_rejections_agree_with_the_c_loader(tmp_path, text, c_code):
    with pytest.raises(AvbdError) as e:
        sceneio.parse_scene(text)
    assert e.value.code == avbd._lib.AVBD_E_INVALID
    p = tmp_path / "bad.scene"
    p.write_text(text)
    with pytest.raises(ValueError, match=f"failed: {c_code}"):
        pyoracle.Oracle(str(p))


def test_frame_format_is_printf_17g():
    o = pyoracle.Oracle(os.path.join(ROOT, "tests", "golden", "scenes", "tumble.scene"),
                        schedule=pyoracle.SCHED_COLOR)
    o.step(40)
    # bodies with the numbers where %.17g has corner cases: signed zero,
    # subnormals, exponent forms, integers past 2^53, the infinities
    o.set_body(0, pos=[-0.0, 2.0 ** -1074, 1e-7], quat=[1e21, 9007199254740993.0, -1e-300, 0.5],
               vel=[math.inf, -math.inf, 123456789.0], angvel=[1e16, 1e17, -2.5e-5])
    st = o.state()
    state = dict(pos=st[:, 0:3], quat=st[:, 3:7], vel=st[:, 7:10], angvel=st[:, 10:13])
    text = sceneio.format_frame(40, state)
    assert text == o.format_state(40)
    back = sceneio.parse_trajectory(text)[40]
    assert np.array_equal(back, st[:, :13])
    assert math.copysign(1, back[0, 0]) < 0


# --- the JavaScript reader and dump (node)
NODE = shutil.which("node")
needs_node = pytest.mark.skipif(NODE is None, reason="node is not installed")

TRICKY = [0.5, -0.0, 0.0, 1.0, 0.1, 1.0 / 3.0, 2.0 ** -1074, 2.2250738585072014e-308, 1e-7, 1e-5, 0.0001,
          0.00012345678901234567, 123456.789, 1e16, 1e17, 12345678901234567.0, 1e21, 1e22, -1e300,
          9007199254740993.0, 131073 / 2 ** 18, 2 ** -20, 0.25881904510252074, 99999999999999999.0, 5e-324,
          1.7976931348623157e308, 4.35, 0.3]


@needs_node
def test_js_g17_is_printf(tmp_path):
    # includes 131073 / 2^18 = 0.500003814697265625: an exact tie at the 17th
    # digit, which %.17g rounds to even and Number's formatters do not
    p = tmp_path / "v.json"
    p.write_text(json.dumps(["-0" if (x == 0 and math.copysign(1, x) < 0) else x for x in TRICKY] +
                            ["inf", "-inf"]))
    r = subprocess.run([NODE, "tests/js/run_scene.mjs", "format", str(p)], capture_output=True, text=True,
                       cwd=ROOT, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split("\n")[:-1] == ["%.17g" % x for x in TRICKY] + ["inf", "-inf"]


@needs_node
@pytest.mark.parametrize("path", SCENES, ids=NAMES)
def test_js_reader_and_dump_match_the_c_loader(path):
    r = subprocess.run([NODE, "tests/js/run_scene.mjs", "dump0", path], capture_output=True, text=True,
                       cwd=ROOT, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout == pyoracle.Oracle(path).format_state(0)


# --- against the reference's own reader (tests/golden/make_scene_vectors.mjs ran
# test/parity/scene_loader.mjs on every fixture and kept what it built)
VECTORS = json.load(open(os.path.join(ROOT, "tests", "golden", "scenes", "reference_loader.json")))


@pytest.mark.parametrize("path", SCENES, ids=NAMES)
def test_readers_build_what_the_reference_reader_builds(path):
    want = VECTORS[NAMES[SCENES.index(path)]]
    wb = np.array(want["bodies"], dtype=np.float64)
    # the package's reader
    scene = sceneio.read_scene(path)
    got = np.array([b["size"] + [b["density"], b["friction"]] + b["pos"] + b["quat"] + b["vel"] + b["angvel"]
                    for b in scene["bodies"]])
    assert np.array_equal(got, np.delete(wb, 5, axis=1))  # column 5 is the mass
    forces = []
    for j in scene["joints"]:
        if j.get("type") == "spring":
            forces.append(["spring", j["a"], j["b"]] + j["rA"] + j["rB"] + [j["stiffness"], j["rest"]])
        else:
            enc = lambda x: "inf" if x == math.inf else ("-inf" if x == -math.inf else x)  # noqa: E731
            forces.append(["joint", j["a"], j["b"]] + j["rA"] + j["rB"] + [enc(j["kLin"]), enc(j["kAng"]),
                                                                        enc(j["fracture"])])
    # the oracle's C reader: mass, pose and velocities of every body, and each
    # spring's rest as it resolves it (a rest below zero becomes the anchors'
    # distance when the spring is made, spring.js:79-84: the reference reader's
    # vectors hold that resolved value)
    o = pyoracle.Oracle(path)
    for k, f in enumerate(forces):
        if f[0] == "spring" and f[-1] < 0:
            f[-1] = float(o.spring(k)[4])
    assert forces == want["forces"]
    st = o.state()
    assert np.array_equal(o.mass()[:, 0], wb[:, 5])
    assert np.array_equal(st[:, :13], wb[:, 6:19])
