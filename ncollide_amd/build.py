"""Builds libncx.so (HIP, gfx950) in-tree with hipcc.

Each source compiles to its own object in parallel (the device code of one file never calls
another file's device functions), then one link makes the shared library.

Flags:
  -ffp-contract=off        no FMA contraction: the reference (Rust) never fuses a*b+c,
                           and every f32 result is compared bit-for-bit with the oracle.
  -packed-fp32-ops off     packed FP32 instructions are not generated.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libncx.so")
OBJ = os.path.join(CSRC, "_build")
SOURCES = ["ncx_core.hip", "ncx_raycast.hip", "ncx_broadphase.hip", "ncx_narrow.hip", "ncx_mlp.hip",
           "ncx_world.hip", "ncx_toi.hip", "ncx_composite.hip", "ncx_manifold.hip", "ncx_hull.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def _compile(src, extra, verbose):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + list(extra) + ["-c", "-o", obj, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=CSRC)
    return obj


def build(verbose=False, extra=()):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, extra, verbose), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=CSRC)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
