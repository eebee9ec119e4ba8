"""Build the native libraries in-tree:
  build/libhyqmom_hip.so      the device step, batch kernels and warm-start network (hipcc, gfx950)
  build/libhyqmom_opcount.so  host build of the cell numerics with FLOP counting (g++)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
         # no packed FP32 instructions: some of their forms have produced wrong results on
         # this hardware without any error
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wno-unknown-pragmas"]


def _run(cmd, tmp, target, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("%s failed" % cmd[0])
    os.replace(tmp, target)
    return target


def build(verbose=True):
    out = os.path.join(HERE, "build")
    os.makedirs(out, exist_ok=True)
    target = os.path.join(out, "libhyqmom_hip.so")
    srcs = [os.path.join(HERE, "csrc", f) for f in ("hyq_kernels.hip", "hyq_mlp.hip")]
    _run([HIPCC] + FLAGS + ["-o", target + ".tmp"] + srcs, target + ".tmp", target, verbose)
    oc = os.path.join(out, "libhyqmom_opcount.so")
    _run([CXX] + HOST_FLAGS + ["-o", oc + ".tmp", os.path.join(HERE, "csrc", "hyq_opcount.cpp")],
         oc + ".tmp", oc, verbose)
    return target


if __name__ == "__main__":
    build()
