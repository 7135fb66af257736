"""Build recipe for the CPU oracle (test infrastructure only).

Compiles oracle/bvh_oracle.c into oracle/liboracle.so with gcc:
-O3, strict IEEE evaluation (-ffp-contract=off, no fast-math), OpenMP for the
multi-core baseline variant.

The reference itself (distBVH) is not built here: its path needs Kokkos,
DARMA/vt and spdlog, none of which exist in this image.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "bvh_oracle.c")
OUT = os.path.join(HERE, "liboracle.so")


def build(verbose=False):
    cmd = [
        "gcc", "-O3", "-std=gnu11", "-fPIC", "-shared", "-fopenmp",
        "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-function",
        SRC, "-o", OUT + ".tmp", "-lm",
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
