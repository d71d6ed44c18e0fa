"""Builds libavbd_hip.so for gfx950 from csrc/ (hipcc, in-tree)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "avbd_api.hip")
OUT = os.path.join(HERE, "libavbd_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # the exact mode relies on IEEE operation-by-operation rounding
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


ADDON_SRC = os.path.join(HERE, "js", "avbd_addon.c")
ADDON_OUT = os.path.join(HERE, "js", "avbd.node")
NODE_INCLUDE = "/usr/include/node"


def build_addon():
    """Node-API addon for the JavaScript host (needs the Node headers)."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        return None
    cmd = ["gcc", "-O2", "-std=c11", "-shared", "-fPIC", "-Wall", "-I", NODE_INCLUDE,
           "-o", ADDON_OUT + ".tmp", ADDON_SRC, "-L", HERE, "-l:libavbd_hip.so",
           "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("addon build failed")
    os.replace(ADDON_OUT + ".tmp", ADDON_OUT)
    return ADDON_OUT


def build(verbose=False):
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), "-o", OUT + ".tmp", SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(OUT + ".tmp", OUT)
    build_addon()
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
