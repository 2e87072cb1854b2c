"""Build libsphhip.so for gfx950 with hipcc (in-tree, next to the package).

Flags:
  --offload-arch=gfx950          MI355X (CDNA4)
  -ffp-contract=off              no FMA contraction: the FP64 mode reproduces
                                 the reference's per-operation rounding
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                 no packed FP32 instructions
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libsphhip.so")

SOURCES = ["sph_ctx.hip", "sph_nnps.hip", "sph_ops.hip", "sph_rate_xsph.hip", "sph_pairs.hip", "sph_batch.hip",
           "sph_staged.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def sources():
    return [os.path.join(HERE, s) for s in SOURCES]


def build(verbose=False):
    cmd = [HIPCC] + FLAGS + ["-I", INCLUDE, "-I", HERE] + sources() + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
