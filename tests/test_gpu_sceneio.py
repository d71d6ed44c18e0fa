"""`.scene` files through the GPU step, dumped as reference_main.cpp dumps
them (test/parity/reference_main.cpp:142-183: load, step, print %.17g state
after every step), against the same dump from the oracle, which reads the
file with its C restatement of the reference loader. Exact precision: the two
dumps must be the same text for all 120 steps of every fixture (the
reference's parity scene set, test/parity/gen_scenes.mjs:71-261)."""
import glob
import os
import shutil
import subprocess

import pytest

import avbd
import pyoracle
from avbd import sceneio

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCENES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "scenes", "*.scene")))
NAMES = [os.path.splitext(os.path.basename(p))[0] for p in SCENES]
STEPS = 120
NODE = shutil.which("node")


def oracle_dump(path, steps):
    o = pyoracle.Oracle(path, schedule=pyoracle.SCHED_COLOR)
    out = []
    for s in range(1, steps + 1):
        o.step(1)
        out.append(o.format_state(s))
    return "".join(out)


@pytest.mark.parametrize("path", SCENES, ids=NAMES)
def test_scene_dump_equals_the_oracle_dump(path):
    got = sceneio.run_scene(path, steps=STEPS, precision=avbd.PREC_EXACT)
    want = oracle_dump(path, STEPS)
    if got != want:
        g, w = sceneio.parse_trajectory(got), sceneio.parse_trajectory(want)
        first = next(s for s in range(1, STEPS + 1) if not (g[s] == w[s]).all())
        pytest.fail(f"first differing step {first}: max |diff| {abs(g[first] - w[first]).max():.3e}")


def test_dump_arguments_follow_reference_main(tmp_path):
    # the positional parameters of reference_main.cpp:159-166, not defaults
    path = os.path.join(ROOT, "tests", "golden", "scenes", "tumble.scene")
    out = tmp_path / "dump.txt"
    with open(out, "w") as f:
        assert sceneio.run_scene(path, 30, 6, 1 / 120, -9.81, 0.9, 2e4, 50.0, 0.99,
                                 precision=avbd.PREC_EXACT, out=f) is None
    params = dict(iterations=6, dt=1 / 120, gravity=-9.81, alpha=0.9, betaLin=2e4, betaAng=50.0, gamma=0.99)
    o = pyoracle.Oracle(path, params, schedule=pyoracle.SCHED_COLOR)
    want = []
    for s in range(1, 31):
        o.step(1)
        want.append(o.format_state(s))
    assert out.read_text() == "".join(want)


@pytest.mark.skipif(NODE is None, reason="node is not installed")
@pytest.mark.parametrize("name", ["rope_angular", "tumble"])
def test_js_host_scene_dump_equals_the_oracle_dump(name):
    path = os.path.join(ROOT, "tests", "golden", "scenes", name + ".scene")
    r = subprocess.run([NODE, "tests/js/run_scene.mjs", "run", path, "60", "exact"], capture_output=True,
                       text=True, cwd=ROOT, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout == oracle_dump(path, 60)
