"""The f64 oracle reproduces the reference CPU engine bit for bit.

Fixtures under tests/golden/ were written by the reference engine itself
(tests/golden/make_golden.mjs): per-step SHA-256 of every body's state and of
the force list, the colouring used under the colour schedule, and exact states
at sample steps. Among them: post-stabilisation (poststab_color), the
body-frame inertia (bodyinertia_color, rotatedInertia false) and bodies removed
and added mid-run (edits_color). Reference: src/physics/solver.js:261-493 and
everything it calls.
"""
import glob
import hashlib
import json
import os

import numpy as np
import pytest

import pyoracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
# trajectory fixtures (pack_arrays.json holds scene-graph extractions instead)
FILES = sorted(f for f in glob.glob(os.path.join(GOLDEN, "*.json")) if not os.path.basename(f).startswith("pack_"))


def _state13(o):
    s = o.state()[:, :13].copy()
    s += 0.0  # fold -0 to +0 (the fixtures do the same)
    return s


def _sha(arr):
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()


def _force_words(o):
    kab, _ = o.forces()
    words = []
    for i, (kind, a, b, nc) in enumerate(kab):
        # the fixtures write 0 for a manifold and 1 for any other force
        words += [min(int(kind), 1), int(a) & 0xFFFFFFFF, int(b), int(nc)]
        if kind == 0:
            words += [int(f) for f in o.manifold(i)["features"]]
    return np.array(words, dtype=np.uint32)


def _run(doc):
    sched = pyoracle.SCHED_REVERSE if doc["schedule"] == "reverse" else pyoracle.SCHED_COLOR
    order = 1 if doc["schedule"] == "prio" else 0
    return pyoracle.Oracle(doc["scene"], doc["params"], schedule=sched, color_order=order)


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f)[:-5] for f in FILES])
def test_bit_exact_against_reference(path):
    doc = json.load(open(path))
    o = _run(doc)
    for s, rec in enumerate(doc["perStep"], start=1):
        o.step(1)
        st = _state13(o)
        if str(s) in doc["samples"]:
            ref = np.array(doc["samples"][str(s)]).reshape(-1, 13)
            diff = np.abs(st - ref).max()
            assert np.array_equal(st, ref), f"step {s}: max |diff| {diff:.3e}"
        fw = _force_words(o)
        if "forceList" in rec:
            assert fw.tolist() == rec["forceList"], f"force list differs at step {s}"
        assert len(o.forces()[0]) == rec["nforces"], f"force count differs at step {s}"
        assert _sha(fw) == rec["forces"], f"force list hash differs at step {s}"
        if "colorOf" in rec:
            col, k = o.last_coloring()
            assert col.tolist() == rec["colorOf"], f"colouring differs at step {s}"
            assert k == rec["numColors"]
        assert _sha(st) == rec["state"], f"state hash differs at step {s}"
        apply_edits(o, doc, s)


def apply_edits(o, doc, s):
    """The topology edits the fixture made after step s: bodies removed by
    their index at that moment (Solver.removeBodies, solver.js:261-300), then
    bodies appended (new Rigid on the running solver)."""
    for e in doc.get("edits", []):
        if e["step"] != s:
            continue
        if e.get("remove"):
            mask = np.zeros(o.n, dtype=np.int32)
            mask[e["remove"]] = 1
            o.remove_bodies(mask)
        if e.get("add"):
            o.add_bodies(e["add"])
