"""Build recipe for the FP64 CPU oracle (test infrastructure only).

Compiles ``oracle/sphref.c`` into ``oracle/libsphref.so`` with gcc,
``-ffp-contract=off`` (no FMA, the expression order of the reference's
baseline x86-64 Cython build) and OpenMP (for the multi-core baseline).

The reference itself (Python 2 + Cython 0.x, PySPH 0.9beta) is not built:
its Cython sources need a Python 2 toolchain and removed numpy APIs, so
there is no ``oracle/_ref`` build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "sphref.c")
LIB = os.path.join(HERE, "libsphref.so")

CFLAGS = ["-O3", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off",
          "-fno-fast-math", "-fopenmp", "-Wall", "-Wno-unknown-pragmas"]


def build(verbose=False):
    cmd = ["gcc"] + CFLAGS + ["-o", LIB + ".tmp", SRC, "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
