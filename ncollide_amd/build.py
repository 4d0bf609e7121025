"""Builds libncx.so (HIP, gfx950) in-tree with hipcc.

Flags:
  -ffp-contract=off        no FMA contraction: the reference (Rust) never fuses a*b+c,
                           and every f32 result is compared bit-for-bit with the oracle.
  -packed-fp32-ops off     packed FP32 instructions are not generated.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libncx.so")
SOURCES = ["ncx_core.hip", "ncx_raycast.hip", "ncx_broadphase.hip", "ncx_narrow.hip", "ncx_mlp.hip",
           "ncx_world.hip", "ncx_hull.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(verbose=False, extra=()):
    srcs = sources()
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + list(extra) + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=CSRC)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
