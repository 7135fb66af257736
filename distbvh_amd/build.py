"""Build recipe of the gfx950 library distbvh_amd/libdistbvh.so.

hipcc cross-compiles for gfx950 without a GPU.  Flags:
  -ffp-contract=off   no FMA contraction: fp64 slabs bit-exact with the CPU path
  -packed-fp32-ops    packed fp32 instructions disabled (target feature off)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["sort.hip", "build.hip", "query.hip", "split.hip", "primitive.hip", "sweep.hip", "wire.hip", "capi.hip"]
LIB = os.path.join(HERE, "libdistbvh.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def build(verbose=False, extra=()):
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        cmd = [HIPCC] + [f for f in FLAGS if f != "-shared"] + list(extra) + [
            "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        objs.append(obj)
    rc = [p.wait() for p in procs]
    if any(rc):
        raise RuntimeError("hipcc failed for %s" % [s for s, r in zip(SOURCES, rc) if r])
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
