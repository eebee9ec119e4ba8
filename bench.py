"""Benchmark: cell-steps/s of the HyQMOM 3D step at 120x120x60 (crossing jets, Ma=1.5,
Kn=0.01, CFL=0.7; --steps steps from the initial state, 200 by default) on one MI355X,
through the C ABI of libhyqmom_hip.so.

Prints one JSON line.  `value` = Nx*Ny*Nz*steps / wall time of the timed steps of the
headline mode (default: mixed precision with the learned warm start), state resident in
HBM, no host copies inside the timed region.  The same workload in plain FP64 is timed
beside it ("fp64"), and the headline mode's one-step error against FP64 from the same
state is reported ("accuracy").

Breakdowns: the 200-step trajectory of this workload passes through near-vacuum cells with
zero variance and runaway velocity, where the reference's z eigenvalue solve throws
(non-finite Jacobian) and this library returns HYQ_ENONFINITE.  Whether a run meets such a
cell within 200 steps is decided by rounding: perturbing the density of 100 cells of the
initial state by one ulp moves the breakdown between step ~90 and beyond 200.  Each timed
run therefore starts from the exact initial state and, if that trajectory breaks down,
from the seeded one-ulp perturbations 1, 2, ... until one completes the steps; the failed
attempts are reported ("initial_state") and not timed.

Roofline: per kernel group (pass 1, update, pass 2: the launches of each phase) the average
duration from HIP events on the context stream over the timed steps; pass 1 and pass 2 are
priced as FP64 vector (VALU) work (algorithmic FLOPs counted on the states of the run by the
op-counting host build of the same cell code) against MI355X's 78.6 TFLOP/s FP64 vector
peak, the update as HBM traffic (algorithmic bytes per cell) against 8 TB/s.  The
"roofline" object is the group with the largest share of the step; its "traffic" is the
HBM bytes per step that rocprofv3's FETCH_SIZE and WRITE_SIZE counters measured for the
same group on the same workload and step count (profiles/hbm_traffic.json, made by the
commands recorded there), or null where no such measurement exists.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hyqmom.jl_amd"))

TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
GRID = (120, 120, 60)


def update_bytes(precision):
    """Algorithmic HBM bytes per cell of the x/y/z update: read the projected state (35
    FP64), the three flux blocks (3x35) and six wave speeds, write the new state."""
    fb = 8 if precision == "fp64" else 4
    return 35 * 8 + 3 * 35 * fb + 6 * fb + 35 * 8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="mixed_learned",
                    choices=["fp64", "mixed", "mixed_newton", "mixed_learned"])
    ap.add_argument("--model", default=None, help="warm-start network blob (mixed_learned)")
    ap.add_argument("--grid", default="%d,%d,%d" % GRID)
    ap.add_argument("--compare-fp64", type=int, default=1)
    ap.add_argument("--compare-newton", type=int, default=1,
                    help="with mixed_learned: also time the same step without the network")
    ap.add_argument("--compare-h128", type=int, default=1,
                    help="with mixed_learned: also time the step with the width-128 network")
    ap.add_argument("--count-flops", type=int, default=1)
    ap.add_argument("--accuracy", type=int, default=1)
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="run one 1 GiB device copy first (HBM counter calibration)")
    return ap.parse_args()


def jets_params(N, precision, model=None):
    import hyqmom_amd as H
    bg, jets = H.crossing_jets(1.5)
    p = dict(Nx=N[0], Ny=N[1], Nz=N[2], tmax=1e9, Kn=0.01, Ma=1.5, flag2D=0, CFL=0.7, Nmom=35,
             nnmax=10 ** 9, dtmax=1e-2, rhol=1.0, rhor=0.1, T=1.0, r110=0.0, r101=0.0, r011=0.0,
             symmetry_check_interval=0, homogeneous_z=False, debug_output=False,
             use_custom_ic=True, ic_background=bg, ic_jets=jets, precision=precision)
    if model:
        p["warmstart_model"] = model
    return p


def perturbed_state(M0, seed):
    """The initial state (seed 0) or the seeded one-ulp perturbation of the density of 100
    cells (seed > 0)."""
    import numpy as np
    if seed == 0:
        return M0
    M = np.array(M0, order="F")
    rng = np.random.default_rng(seed)
    idx = tuple(rng.integers(0, n, 100) for n in M.shape[:3])
    M[idx + (0,)] = np.nextafter(M[idx + (0,)], np.inf)
    return M


class Breakdown(Exception):
    def __init__(self, step):
        super().__init__("non-finite at step %d" % step)
        self.step = step


def timed_run(N, precision, M0, steps, warmup, model=None):
    """Wall time and per-kernel event times of `steps` steps from M0; Breakdown if the
    trajectory goes non-finite."""
    import torch
    import hyqmom_amd as H
    from hyqmom_amd._lib import HyqError
    s = H.Solver(jets_params(N, precision, model))
    s.set_state(M0)
    s.run(1e9, warmup)
    s.set_state(M0)
    s.t, s.nn = 0.0, 0
    s.reset_kernel_times()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    try:
        s.run(1e9, steps)
    except HyqError:
        nn = s.nn
        s.close()
        raise Breakdown(nn)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    p1, up, p2, nst = s.kernel_times()
    assert nst == steps and s.nn == steps, (nst, s.nn)
    attached = s._mlp is not None
    fl = s.flags()  # pass-1 flags of the last step
    d1, d2 = s.deferred_cells()
    why = s.fast_path_counts() if precision in ("mixed_newton", "mixed_learned") else None
    s.close()
    from hyqmom_amd._lib import newton_iterations
    return {"seconds": el, "pass1_ms": p1, "update_ms": up, "pass2_ms": p2, "network": attached,
            "last_step": {"fallback_cells_frac": float(((fl >> 12) & 1).mean()),
                          "iterations_per_cell": float(newton_iterations(fl).mean()),
                          "deferred_cells_frac": [d1 / fl.size, d2 / fl.size],
                          "fast_path_counts": why}}


def surviving_run(N, precision, M0, steps, warmup, model=None, tries=24):
    """timed_run from the initial state, else from the first seeded perturbation whose
    trajectory completes the steps.  Returns (run, start state, record of the attempts)."""
    failed = []
    for seed in range(tries):
        M = perturbed_state(M0, seed)
        try:
            run = timed_run(N, precision, M, steps, warmup, model)
        except Breakdown as b:
            failed.append({"seed": seed, "breakdown_step": b.step})
            continue
        return run, M, {"seed": seed, "failed_attempts": failed}
    raise RuntimeError("%s: no trajectory completed %d steps in %d attempts: %s"
                       % (precision, steps, tries, failed))


def replay_states(N, precision, M0, steps, nsample, model=None):
    """States at nsample evenly spaced steps of the same run (untimed replay)."""
    import hyqmom_amd as H
    s = H.Solver(jets_params(N, precision, model))
    s.set_state(M0)
    marks = sorted(set(int(round(k * (steps - 1) / max(1, nsample - 1))) for k in range(nsample)))
    out = {}
    for m in marks:
        if m > s.nn:
            s.run(1e9, m)
        out[m] = s.get_state()
    s.close()
    return out


def count_flops(states, precision, stride=211):
    """Mean pass-1 / pass-2 FP64 FLOPs per cell over the sampled states."""
    from hyqmom_amd import opcount
    newton = precision in ("mixed_newton", "mixed_learned")
    a = b = 0.0
    for M in states.values():
        f1, f2 = opcount.state_flops_per_cell(M, stride=stride, Ma=1.5, dt=1e-3, Kn=0.01, newton=newton)
        a += f1
        b += f2
    return a / len(states), b / len(states)


def fields(M):
    r = M[..., 0]
    return {"rho": r, "u": M[..., 1] / r, "v": M[..., 5] / r, "w": M[..., 15] / r,
            "E": 0.5 * (M[..., 2] + M[..., 9] + M[..., 19])}


def one_step(N, precision, M, model=None):
    """One step from M.  A non-finite z Jacobian (HYQ_ENONFINITE) still completes the step on
    the device; the state is returned with whatever the step left in that cell."""
    import hyqmom_amd as H
    from hyqmom_amd._lib import HyqError
    s = H.Solver(jets_params(N, precision, model))
    s.set_state(M)
    try:
        s.step()
    except HyqError as e:
        if "HYQ_ENONFINITE" not in str(e):
            raise
    out = s.get_state()
    s.close()
    return out


def accuracy(N, precision, M, model=None):
    """One step of `precision` and of FP64 from the same FP64-trajectory state.  Cells whose
    FP64 result itself moves by more than 1e-8 (field-scaled) when the nonzero moments of the
    input are perturbed by one ulp, or becomes non-finite, are ill-conditioned (realizability / eigenvalue branch
    knife edges) and are counted, not compared; the max relative error (field-scaled) is
    over the other cells."""
    import numpy as np
    ref = one_step(N, "fp64", M)
    got = one_step(N, precision, M, model)
    # the perturbed step can itself meet a non-finite z Jacobian, which makes its dt and so
    # every cell non-finite: then another perturbation is drawn, and if none gives a finite
    # step, no cell is classed ill-conditioned
    pert, used = None, None
    for seed in range(7, 15):
        rng = np.random.default_rng(seed)
        Mp = np.asfortranarray(M.copy())
        # one ulp on half of the nonzero moments; zero moments (odd moments of symmetric
        # states) stay zero, a subnormal there being no rounding-level change of the state
        flip = (rng.random(M.shape) < 0.5) & (M != 0)
        Mp[flip] = np.nextafter(Mp[flip], np.inf)
        p = one_step(N, "fp64", Mp)
        if np.isfinite(p[..., 0]).mean() > 0.99:
            pert, used = p, seed
            break
    fr, fg = fields(ref), fields(got)
    bad = np.zeros(M.shape[:3], bool)  # non-finite in the FP64 step itself
    for k in fr:
        bad |= ~np.isfinite(fr[k])
    ill = np.zeros(M.shape[:3], bool)
    if pert is not None:
        fp = fields(pert)
        for k in fr:
            sc = np.abs(fr[k][~bad]).max()
            with np.errstate(invalid="ignore"):
                ill |= ~bad & ~(np.abs(fp[k] - fr[k]) / sc <= 1e-8)
    keep = ~bad & ~ill
    out = {"ill_conditioned_cells": int(ill.sum()), "nonfinite_cells": int(bad.sum()),
           "perturbation_seed": used, "mode_step_finite": float(np.isfinite(got[..., 0]).mean())}
    for k in fr:
        if not keep.any():
            out[k] = None
            continue
        sc = np.abs(fr[k][keep]).max()
        e = np.abs(fg[k] - fr[k]) / sc
        out[k] = float(e[keep].max())
    vals = [out[k] for k in fr if out[k] is not None]
    out["max"] = max(vals) if vals else None
    return out


def oracle_rate(M, lo, n, nthreads, seconds, max_steps):
    """cell-steps/s of the oracle on the n^3 block of cells lo..lo+n-1 in x, y and z."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    sub = np.asfortranarray(M[lo:lo + n, lo:lo + n, lo:lo + n, :])
    t0 = time.time()
    steps = 0
    while time.time() - t0 < seconds and steps < max_steps:
        sub, _, _, _, _ = O.run(sub, n, n, n, 0.01, 1.5, 0, 0.7, 1e-2, 1 / 120, 1 / 120, 1 / 60, True,
                                1e9, 1, nthreads=nthreads)
        steps += 1
    return n ** 3 * steps / (time.time() - t0), steps


def cpu_baseline(M, step, seconds=10.0):
    """The oracle (FP64 CPU restatement of the reference step, no MPI: one rank) on this
    host's cores, on a bounded sample of the same workload: steps of the 16x16x16 block
    around a jet's centre (cells 22..37 in x, y and z), cut from the FP64 state at the given
    step, with one OpenMP thread per core; and steps of the 8x8x8 block at its centre (cells
    26..33) on one thread."""
    nth = os.cpu_count() or 1
    v, steps = oracle_rate(M, 22, 16, nth, seconds, 50)
    v1, steps1 = oracle_rate(M, 26, 8, 1, seconds / 2, 50)
    return {"value": v, "unit": "cell-steps/s", "cores": nth, "kind": "port", "mpi_ranks": 1,
            "cpu_model": host_cpu_model(),
            "sample": "%d consecutive steps of the 16^3 block around a jet centre of the 120x120x60 "
                      "crossing-jets FP64 state at step %d, %d OpenMP threads" % (steps, step, nth),
            "one_thread": {"value": v1, "unit": "cell-steps/s",
                           "sample": "%d consecutive steps of the 8^3 block at the same jet centre, "
                                     "1 thread" % steps1}}


def traffic_key(precision, N3, steps):
    return "%s %dx%dx%d %d steps" % ((precision,) + tuple(N3) + (steps,))


def measured_traffic(precision, N3, steps, group):
    """HBM bytes per step of a kernel group measured by rocprofv3 on this workload, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f)["runs"].get(traffic_key(precision, N3, steps))
    except (OSError, ValueError, KeyError):
        return None
    return None if rec is None else rec.get(group)


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def kernel_report(run, flops, precision, N):
    p1f, p2f = flops if flops else (None, None)
    ks = {}
    for name, ms, fl in (("pass1", run["pass1_ms"], p1f), ("pass2", run["pass2_ms"], p2f)):
        d = {"ms": ms, "bound": "valu", "unit": "TFLOP/s", "peak": PEAK_FP64_TFLOPS}
        if fl is not None:
            d["flops_per_cell"] = fl
            d["achieved"] = fl * N / (ms * 1e-3) / 1e12
            d["frac"] = d["achieved"] / PEAK_FP64_TFLOPS
        ks[name] = d
    ub = update_bytes(precision)
    ach = ub * N / (run["update_ms"] * 1e-3) / 1e9
    ks["update"] = {"ms": run["update_ms"], "bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS,
                    "bytes_per_cell": ub, "achieved": ach, "frac": ach / PEAK_HBM_GBS}
    return ks


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (the domain is "
              "one device; see the sharding notes in the design document)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import hyqmom_amd as H
    N3 = tuple(int(x) for x in a.grid.split(","))
    N = N3[0] * N3[1] * N3[2]
    if a.copy_calibration:
        x = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        del x, y
    M0 = np.asfortranarray(H.initial_state(jets_params(N3, "fp64")))

    run, Mrun, start = surviving_run(N3, a.precision, M0, a.steps, a.warmup, a.model)
    value = N * a.steps / run["seconds"]
    flops = None
    states = None
    if a.count_flops or a.accuracy:
        states = replay_states(N3, a.precision, Mrun, a.steps, 6, a.model)
    if a.count_flops:
        flops = count_flops(states, a.precision)
    ks = kernel_report(run, flops, a.precision, N)
    dom = max(ks, key=lambda k: ks[k]["ms"])
    d = ks[dom]
    traffic = measured_traffic(a.precision, N3, a.steps, dom)
    roof = {"kernel": dom, "bound": d["bound"], "achieved": d.get("achieved"), "peak": d["peak"],
            "unit": d["unit"], "frac": d.get("frac"), "traffic": traffic}
    out = {
        "metric": "cell-steps/s at 120\u00b2\u00d760; mixed+learned vs plain FP64 at max rel err \u22641e-6",
        "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": run["seconds"] * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f64+f32" + ("+bf16" if a.precision == "mixed_learned" else ""),
        "data": "synthetic crossing-jets initial condition",
        "initial_state": start,
        "config": {"workload": "crossing jets %dx%dx%d, Ma=1.5, Kn=0.01, CFL=0.7, %d steps, %s"
                               % (N3 + (a.steps, a.precision)),
                   "cells": N, "precision": a.precision,
                   "warmstart_network": run["network"] if a.precision == "mixed_learned" else None},
        "fast_wave_speeds_last_step": run["last_step"] if a.precision in ("mixed_newton", "mixed_learned") else None,
        "kernels": ks,
        "roofline": roof,
    }
    # the FP64 trajectory, sampled: FLOPs of the FP64 kernels, the accuracy check's and the
    # CPU baseline's starting state (mid-run, when the jets have crossed)
    fstates, fstart = (states, start) if a.precision == "fp64" else (None, None)
    f = None
    if a.compare_fp64 and a.precision != "fp64":
        f, Mf, fstart = surviving_run(N3, "fp64", M0, a.steps, a.warmup)
    if fstates is None and ((a.compare_fp64 and a.count_flops) or a.accuracy or a.cpu_baseline):
        if fstart is None:
            _, Mf, fstart = surviving_run(N3, "fp64", M0, a.steps, 0)
        fstates = replay_states(N3, "fp64", Mf, a.steps, 6)
    mid = sorted(fstates)[len(fstates) // 2] if fstates else 0
    Mmid = fstates[mid] if fstates else M0
    if f is not None:
        fv = N * a.steps / f["seconds"]
        fflops = count_flops(fstates, "fp64") if a.count_flops else None
        out["fp64"] = {"value": fv, "ms_per_step": f["seconds"] * 1e3 / a.steps,
                       "kernels": kernel_report(f, fflops, "fp64", N), "initial_state": fstart}
        out["speedup_vs_fp64"] = value / fv
    if a.compare_newton and a.precision == "mixed_learned":
        w, _, wstart = surviving_run(N3, "mixed_newton", M0, a.steps, a.warmup)
        wv = N * a.steps / w["seconds"]
        out["mixed_newton"] = {"value": wv, "ms_per_step": w["seconds"] * 1e3 / a.steps,
                               "kernels": {k: w[k] for k in ("pass1_ms", "update_ms", "pass2_ms")},
                               "fast_wave_speeds_last_step": w["last_step"],
                               "initial_state": wstart}
        if "fp64" in out:
            out["mixed_newton"]["speedup_vs_fp64"] = wv / out["fp64"]["value"]
    if a.compare_h128 and a.precision == "mixed_learned" and a.model is None:
        from hyqmom_amd.warmstart import MODEL_H128
        w, _, wstart = surviving_run(N3, "mixed_learned", M0, a.steps, a.warmup, MODEL_H128)
        wv = N * a.steps / w["seconds"]
        out["mixed_learned_h128"] = {"value": wv, "ms_per_step": w["seconds"] * 1e3 / a.steps,
                                     "network": "15-128-128-128-2",
                                     "kernels": {k: w[k] for k in ("pass1_ms", "update_ms", "pass2_ms")},
                                     "fast_wave_speeds_last_step": w["last_step"],
                                     "initial_state": wstart}
        if "fp64" in out:
            out["mixed_learned_h128"]["speedup_vs_fp64"] = wv / out["fp64"]["value"]
    if a.accuracy and a.precision != "fp64":
        out["accuracy"] = dict(accuracy(N3, a.precision, Mmid, a.model), from_step=mid)
    if a.cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(Mmid, mid)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
