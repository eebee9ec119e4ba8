"""Train the warm-start network on plane Jacobians dumped from the plain solver's own runs.

    python -m hyqmom_amd.train_warmstart --out models/warmstart_mlp.f32

Samples: the five plane Jacobians (UV, UW, VU, VW, WU) of every cell of states taken from
FP64 runs of the crossing-jets problem on 20x20x10 (Ma=1, Kn=1, to t=0.01) and on
120x120x60 (Ma=1.5, Kn=0.01), after the realizability projection the step applies before
its wave-speed solve.  Labels are the standardized extreme eigenvalues of the full FP64
solve; planes with complex or non-finite eigenvalues are left out (the step never uses a
warm start there: the hyperbolicity test sends them to the full solve).

The last production-grid state is held out: the report gives the prediction error on it
and the Laguerre iterations / fallback rate of one step from it with the analytic start
("mixed_newton") and with the network ("mixed_learned").
"""
import argparse
import ctypes
import json
import os
import time

import numpy as np

from . import _lib
from . import warmstart as W
from .initial_conditions import crossing_jets
from .runner import Solver, initial_state


def jets_params(N, Ma, Kn, precision="fp64", tmax=1e9):
    bg, jets = crossing_jets(Ma)
    return dict(Nx=N[0], Ny=N[1], Nz=N[2], tmax=tmax, Kn=Kn, Ma=Ma, flag2D=0, CFL=0.7, Nmom=35,
                nnmax=10 ** 9, dtmax=1e-2, rhol=1.0, rhor=0.1, T=1.0, r110=0.0, r101=0.0,
                r011=0.0, symmetry_check_interval=0, homogeneous_z=False, debug_output=False,
                use_custom_ic=True, ic_background=bg, ic_jets=jets, precision=precision)


def collect_states(N, Ma, Kn, steps_wanted, tmax=1e9, log=print):
    """Host states (Fortran (Nx,Ny,Nz,35)) of an FP64 run at the given step numbers."""
    p = jets_params(N, Ma, Kn, "fp64", tmax)
    s = Solver(p)
    s.set_state(initial_state(p))
    out = {}
    last = max(steps_wanted)
    while True:
        if s.nn in steps_wanted:
            out[s.nn] = s.get_state()
        if s.nn >= last or s.t >= tmax:
            break
        s.step()
    s.close()
    log("collected %d states of %s (t=%.4g)" % (len(out), N, s.t))
    return out


def projected_planes(M_host, Ma):
    """Device (X, L, valid) of the realizability-projected states of one host state."""
    import torch
    n = M_host.shape[0] * M_host.shape[1] * M_host.shape[2]
    soa = torch.from_numpy(np.ascontiguousarray(M_host.ravel(order="F").reshape(35, n))).cuda()
    Mr = torch.empty_like(soa)
    torch.cuda.synchronize()
    _lib.check(_lib.lib().hyq_closure_batch(ctypes.c_void_p(soa.data_ptr()), n, float(Ma), 0,
                                            ctypes.c_void_p(Mr.data_ptr()), None, None, None, None),
               "hyq_closure_batch")
    return W.plane_dataset(Mr)


def dataset(states, max_samples, seed, log=print):
    """states: list of (host state, Ma)."""
    import torch
    Xs, Ls = [], []
    for M, Ma in states:
        X, L, V = projected_planes(M, Ma)
        keep = V.bool()
        Xs.append(X[keep])
        Ls.append(L[keep])
        log("  state %s: %d planes, %d valid" % (M.shape[:3], V.numel(), int(keep.sum())))
    X = torch.cat(Xs)
    L = torch.cat(Ls)
    if X.shape[0] > max_samples:
        g = torch.Generator(device=X.device)
        g.manual_seed(seed)
        idx = torch.randperm(X.shape[0], device=X.device, generator=g)[:max_samples]
        X, L = X[idx].contiguous(), L[idx].contiguous()
    return X, L


def newton_report(M_host, N, Ma, Kn, model):
    """One step from M_host with the analytic start and with the network."""
    rep = {}
    for prec in ("mixed_newton", "mixed_learned"):
        p = jets_params(N, Ma, Kn, prec)
        if prec == "mixed_learned":
            p["warmstart_model"] = model
        s = Solver(p)
        s.set_state(M_host)
        s.step()
        s.reset_kernel_times()
        s.set_state(M_host)
        s.t, s.nn = 0.0, 0
        s.step()
        a, b, d, _ = s.kernel_times()
        fl = s.flags()
        it = _lib.newton_iterations(fl)
        rep[prec] = {"newton_iters_per_cell": float(it.mean()),
                     "fallback_cells": int(((fl >> 12) & 1).sum()),
                     "cells": int(fl.size),
                     "pass1_ms": a, "update_ms": b, "pass2_ms": d}
        s.close()
    return rep


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=W.DEFAULT_MODEL)
    ap.add_argument("--report", default=None)
    ap.add_argument("--samples", type=int, default=10_000_000)
    ap.add_argument("--train-steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--c2-steps", default="0,5,10,20,30")
    ap.add_argument("--holdout-step", type=int, default=40)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--hidden", type=int, default=32, choices=list(W.HIDDEN_WIDTHS))
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    import torch
    t0 = time.time()
    c1 = collect_states((20, 20, 10), 1.0, 1.0, set(range(0, 100)), tmax=0.01, log=log)
    want = set(int(x) for x in a.c2_steps.split(",")) | {a.holdout_step}
    c2 = collect_states((120, 120, 60), 1.5, 0.01, want, log=log)
    hold = c2.pop(a.holdout_step)
    log("states collected in %.1fs" % (time.time() - t0))
    X, L = dataset([(M, 1.0) for M in c1.values()] + [(M, 1.5) for M in c2.values()], a.samples,
                   a.seed, log)
    log("training samples: %d" % X.shape[0])
    t1 = time.time()
    model, hist = W.train(X, L, steps=a.train_steps, batch=a.batch, lr=a.lr, lr_final=a.lr * 0.05,
                          seed=a.seed, log=log, hidden=a.hidden)
    torch.cuda.synchronize()
    train_s = time.time() - t1
    model.save(a.out)
    log("saved %s (training %.1fs)" % (a.out, train_s))

    Xh, Lh, Vh = projected_planes(hold, 1.5)
    keep = Vh.bool()
    P = model.infer(Xh[keep]).cpu().numpy()
    T = Lh[keep].cpu().numpy()
    err = np.abs(P - T)
    rep = {"hidden": a.hidden, "samples": int(X.shape[0]), "train_steps": a.train_steps, "batch": a.batch,
           "train_seconds": train_s, "loss_first": hist[0][1], "loss_last": hist[-1][1],
           "holdout_planes": int(keep.sum()),
           "holdout_abs_err": {"median": float(np.median(err)), "p95": float(np.percentile(err, 95)),
                               "p99": float(np.percentile(err, 99)), "max": float(err.max())},
           "holdout_step": a.holdout_step}
    rep["step_from_holdout"] = newton_report(hold, (120, 120, 60), 1.5, 0.01, model)
    log(json.dumps(rep, indent=1))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    return rep


if __name__ == "__main__":
    main()
