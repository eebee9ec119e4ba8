"""Benchmark: cell-steps/s of the HyQMOM 3D step at 120x120x60 (crossing jets, Ma=1.5,
Kn=0.01, CFL=0.7) on one MI355X, through the C ABI of libhyqmom_hip.so.

Prints one JSON line.  `value` = Nx*Ny*Nz*steps / wall time of the timed steps (state
resident in HBM; no host copies inside the timed region).  The roofline object refers to
the dominant kernel (pass 1), timed with HIP events on the context stream.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hyqmom.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

# FP64 operations per cell-step of the pass-1 kernel and per update+pass-2 pair, from the
# oracle's op count on the crossing-jets state; bytes = algorithmic HBM
# traffic per cell.
PASS1_BYTES = 280 + 280 + 840 + 48 + 4        # read M; write Mr, Fx/Fy/Fz, vp, flags
STEP_BYTES_FP64 = 2896


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed", "mixed_newton", "mixed_learned"])
    ap.add_argument("--model", default=None, help="warm-start network blob (mixed_learned)")
    ap.add_argument("--grid", default="120,120,60")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    return ap.parse_args()


def cpu_baseline(M0, params, ncells_sample=4096):
    """The oracle (FP64 CPU restatement) on the host cores: pass 1 + step on a bounded
    sample of the same workload (a 16x16x16 block cut from the 120x120x60 state)."""
    import numpy as np
    import oracle as O
    nth = os.cpu_count() or 1
    sub = np.asfortranarray(M0[52:68, 52:68, 22:38, :])
    t0 = time.time()
    steps = 0
    while time.time() - t0 < 10.0 and steps < 50:
        O.run(sub, 16, 16, 16, params["Kn"], params["Ma"], 0, params["CFL"], params["dtmax"],
              1 / 120, 1 / 120, 1 / 60, True, 1e9, 1, nthreads=nth)
        steps += 1
    el = time.time() - t0
    return {"value": 16 * 16 * 16 * steps / el, "unit": "cell-steps/s", "cores": nth, "kind": "port",
            "sample": "%d steps of a 16^3 block of the 120x120x60 crossing-jets state, %d threads" % (steps, nth)}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU path)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import hyqmom_amd as H
    Nx, Ny, Nz = (int(x) for x in a.grid.split(","))
    bg, jets = H.crossing_jets(1.5)
    params = dict(Nx=Nx, Ny=Ny, Nz=Nz, tmax=1e9, Kn=0.01, Ma=1.5, flag2D=0, CFL=0.7, Nmom=35,
                  nnmax=10**9, dtmax=1e-2, rhol=1.0, rhor=0.1, T=1.0, r110=0.0, r101=0.0, r011=0.0,
                  symmetry_check_interval=0, homogeneous_z=False, debug_output=False,
                  use_custom_ic=True, ic_background=bg, ic_jets=jets, precision=a.precision)
    if a.model:
        params["warmstart_model"] = a.model
    M0 = H.initial_state(params)
    s = H.Solver(params)
    s.set_state(M0)
    s.run(1e9, a.warmup)
    s.reset_kernel_times()
    import ctypes
    from hyqmom_amd import _lib
    t0 = time.perf_counter()
    s.run(1e9, a.warmup + a.steps)
    el = time.perf_counter() - t0
    ms_p1, ms_up, nst = s.kernel_times()
    N = Nx * Ny * Nz
    value = N * a.steps / el
    ach = PASS1_BYTES * N / (ms_p1 * 1e-3) / 1e9
    out = {
        "metric": "cell-steps/s at 120x120x60 (crossing jets, Ma=1.5, Kn=0.01)",
        "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": el * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if a.precision == "fp64" else "f64+f32",
        "data": "synthetic crossing-jets initial condition",
        "config": {"workload": "crossing jets %dx%dx%d, Ma=1.5, Kn=0.01, CFL=0.7, %s" % (Nx, Ny, Nz, a.precision),
                   "cells": N, "precision": a.precision},
        "kernel_ms": {"pass1": ms_p1, "update_pass2": ms_up, "steps_timed": nst},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": 8000.0, "unit": "GB/s",
                     "frac": ach / 8000.0, "traffic": None},
    }
    if a.cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(M0, params)
    s.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
