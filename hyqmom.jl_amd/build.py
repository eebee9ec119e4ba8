"""Build libhyqmom_hip.so in-tree for gfx950 (hipcc)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
         # packed FP32 instructions off (see the environment notes of this repository)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=True):
    out = os.path.join(HERE, "build")
    os.makedirs(out, exist_ok=True)
    target = os.path.join(out, "libhyqmom_hip.so")
    tmp = target + ".tmp"
    srcs = [os.path.join(HERE, "csrc", f) for f in ("hyq_kernels.hip", "hyq_mlp.hip")]
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, target)
    return target


if __name__ == "__main__":
    build()
