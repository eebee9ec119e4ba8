"""Build the native libraries in-tree:
  build/libhyqmom_hip.so      the device step, batch kernels, ensemble and warm-start network
                              (hipcc, gfx950; each source compiled in its own process, then linked)
  build/libhyqmom_opcount.so  host build of the cell numerics with FLOP counting (g++)
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC",
         # no packed FP32 instructions: some of their forms have produced wrong results on
         # this hardware without any error
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         # multiply-add fusion within one source expression only: the same cell code then
         # rounds the same way in every kernel it is inlined into (single run, deferred
         # cells, ensemble), which keeps those paths bitwise identical
         "-ffp-contract=on"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wno-unknown-pragmas"]
# (source, object, extra flags): the warm-start network is built once per hidden width
SOURCES = [("hyq_kernels.hip", "hyq_kernels.o", []), ("hyq_ensemble.hip", "hyq_ensemble.o", []),
           ("hyq_ens_fp64.hip", "hyq_ens_fp64.o", []), ("hyq_ens_mixed.hip", "hyq_ens_mixed.o", []),
           ("hyq_ens_newton.hip", "hyq_ens_newton.o", []),
           ("hyq_mlp.hip", "hyq_mlp_h32.o", ["-DHYQ_MLP_HID=32", "-DHYQ_MLP_NS=mlp_h32"]),
           ("hyq_mlp.hip", "hyq_mlp_h128.o", ["-DHYQ_MLP_HID=128", "-DHYQ_MLP_NS=mlp_h128"]),
           ("hyq_mlp_api.hip", "hyq_mlp_api.o", [])]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("%s failed" % cmd[0])


def build(verbose=True):
    out = os.path.join(HERE, "build")
    os.makedirs(out, exist_ok=True)
    objs = [os.path.join(out, o) for _, o, _ in SOURCES]
    cmds = [[HIPCC] + FLAGS + extra + ["-c", "-o", o, os.path.join(HERE, "csrc", f)]
            for (f, _, extra), o in zip(SOURCES, objs)]
    with ThreadPoolExecutor(len(cmds)) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    target = os.path.join(out, "libhyqmom_hip.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", target + ".tmp"] + objs, verbose)
    os.replace(target + ".tmp", target)
    oc = os.path.join(out, "libhyqmom_opcount.so")
    _run([CXX] + HOST_FLAGS + ["-o", oc + ".tmp", os.path.join(HERE, "csrc", "hyq_opcount.cpp")], verbose)
    os.replace(oc + ".tmp", oc)
    return target


if __name__ == "__main__":
    build()
