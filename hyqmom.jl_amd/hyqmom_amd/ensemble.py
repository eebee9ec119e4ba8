"""Many independent crossing-jets runs advanced together on one device (libhyqmom_hip.so
hyq_ensemble_*): shared SoA planes, one HIP-graph-captured step, per-run CFL dt, and an
autotuner for the block size of each cell kernel.  No counterpart in HyQMOM.jl, which runs
one simulation per process."""
import ctypes
import math
import time

import numpy as np

from . import _lib
from .initial_conditions import crossing_jets
from .runner import Solver, initial_state

C5_GRIDS = [(16, 16, 16), (24, 24, 24), (32, 32, 16), (48, 48, 48), (64, 64, 32)]


def jets_params(N, Ma, Kn, precision="mixed", tmax=1e9, CFL=0.7):
    bg, jets = crossing_jets(Ma)
    return dict(Nx=N[0], Ny=N[1], Nz=N[2], tmax=tmax, Kn=Kn, Ma=Ma, flag2D=0, CFL=CFL, Nmom=35,
                nnmax=10 ** 9, dtmax=1e-2, rhol=1.0, rhor=0.1, T=1.0, r110=0.0, r101=0.0,
                r011=0.0, symmetry_check_interval=0, homogeneous_z=False, debug_output=False,
                use_custom_ic=True, ic_background=bg, ic_jets=jets, precision=precision)


def ensemble_runs(n=64, seed=0x5EED0005, precision="mixed"):
    """The batched-ensemble workload: grids uniform from C5_GRIDS, Ma ~ U[0.5, 2],
    Kn log-uniform in [1e-3, 1]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        N = C5_GRIDS[int(rng.integers(0, len(C5_GRIDS)))]
        Ma = float(rng.uniform(0.5, 2.0))
        Kn = float(10 ** rng.uniform(-3.0, 0.0))
        out.append(jets_params(N, Ma, Kn, precision))
    return out


def _hyq_params(p):
    q = _lib.HyqParams()
    q.Nx, q.Ny, q.Nz = int(p["Nx"]), int(p["Ny"]), int(p["Nz"])
    q.Nmom = 35
    q.flag2D = int(p["flag2D"])
    q.positivity = 1 if p.get("positivity", True) else 0
    q.precision = _lib.PREC[p.get("precision", "fp64")]
    q.tmax, q.Kn, q.Ma, q.CFL, q.dtmax = (float(p[k]) for k in ("tmax", "Kn", "Ma", "CFL", "dtmax"))
    q.xmin, q.xmax = p.get("xmin", -0.5), p.get("xmax", 0.5)
    q.ymin, q.ymax = p.get("ymin", -0.5), p.get("ymax", 0.5)
    q.zmin, q.zmax = p.get("zmin", -0.5), p.get("zmax", 0.5)
    return q


class Ensemble:
    def __init__(self, runs):
        self.runs = list(runs)
        arr = (_lib.HyqParams * len(self.runs))(*[_hyq_params(p) for p in self.runs])
        h = ctypes.c_void_p()
        self._check(_lib.lib().hyq_ensemble_create(arr, len(self.runs), ctypes.byref(h)), "hyq_ensemble_create")
        self.h = h
        self.shapes = [(int(p["Nx"]), int(p["Ny"]), int(p["Nz"]), 35) for p in self.runs]
        self.cells = sum(s[0] * s[1] * s[2] for s in self.shapes)

    @staticmethod
    def _check(rc, what):
        if rc != 0:
            raise _lib.HyqError("%s failed: %s (%s)" % (what, _lib.ERRORS.get(rc, rc),
                                                        _lib.lib().hyq_ensemble_last_error().decode()))

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().hyq_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, r, M):
        flat = np.asfortranarray(np.asarray(M, np.float64)).ravel(order="F")
        self._check(_lib.lib().hyq_ensemble_set_state(self.h, r, flat.ctypes.data_as(_lib.D)), "set_state")

    def set_initial_states(self):
        for r, p in enumerate(self.runs):
            self.set_state(r, initial_state(p))

    def get_state(self, r):
        shp = self.shapes[r]
        flat = np.empty(int(np.prod(shp)))
        self._check(_lib.lib().hyq_ensemble_get_state(self.h, r, flat.ctypes.data_as(_lib.D)), "get_state")
        return flat.reshape(shp, order="F")

    def status(self):
        n = len(self.runs)
        t, dt, nn = np.empty(n), np.empty(n), np.empty(n, np.int64)
        self._check(_lib.lib().hyq_ensemble_status(self.h, t.ctypes.data_as(_lib.D), dt.ctypes.data_as(_lib.D),
                                                   nn.ctypes.data_as(ctypes.POINTER(_lib.I64))), "status")
        return t, dt, nn

    def run(self, nsteps, graph=True, strict=True):
        """Advance every active run by up to nsteps.  A run whose z wave-speed solve meets a
        non-finite Jacobian (where the reference throws) stops there while the others go on;
        with strict=True that raises HyqError (HYQ_ENONFINITE) naming the stopped runs after
        the others have finished their steps, with strict=False it is left to status()."""
        rc = _lib.lib().hyq_ensemble_run(self.h, int(nsteps), 1 if graph else 0)
        if rc == _lib.HYQ_ENONFINITE and not strict:
            return
        self._check(rc, "run")

    def set_blocks(self, blocks):
        b = (ctypes.c_int32 * 3)(*blocks)
        self._check(_lib.lib().hyq_ensemble_set_blocks(self.h, b), "set_blocks")

    def tune(self, trial_steps=2):
        b = (ctypes.c_int32 * 3)()
        ms = np.empty(9)
        self._check(_lib.lib().hyq_ensemble_tune(self.h, int(trial_steps), b, ms.ctypes.data_as(_lib.D)), "tune")
        return list(b), ms.reshape(3, 3)


def compare_sequential(runs, steps=10, tune=True, log=print):
    """Aggregate cell-steps/s of the graph-captured ensemble against the same runs stepped
    one after another, each through its own single-run context.  Only steps a run actually
    takes count (a run that has reached tmax, or gone non-finite where the reference
    throws, stops in both)."""
    import torch
    e = Ensemble(runs)
    e.set_initial_states()
    blocks = None
    if tune:
        blocks, ms = e.tune(2)
        log("autotuned blocks (pass1, update, pass2): %s; trial ms per candidate 64/128/256:\n%s" % (blocks, ms))
    e.run(1, strict=False)  # graph build and first replay
    _, _, nn0 = e.status()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.run(steps, strict=False)
    el_e = time.perf_counter() - t0
    _, _, nn1 = e.status()
    sizes = np.array([s[0] * s[1] * s[2] for s in e.shapes], dtype=np.int64)
    work_e = int((sizes * (nn1 - nn0)).sum())
    e.close()
    solvers = []
    for p in runs:
        s = Solver(p)
        s.set_state(initial_state(p))
        live = True
        try:
            s.step()
        except _lib.HyqError:
            live = False
        solvers.append([s, live, int(p["Nx"]) * int(p["Ny"]) * int(p["Nz"])])
    torch.cuda.synchronize()
    work_s = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        for ent in solvers:
            s, live, n = ent
            if not live or not (s.t < float(s.p.tmax)):
                continue
            try:
                s.step()
                work_s += n
            except _lib.HyqError:
                ent[1] = False
                work_s += n
    el_s = time.perf_counter() - t0
    for ent in solvers:
        ent[0].close()
    out = {"runs": len(runs), "cells": int(sizes.sum()), "steps": steps, "blocks": blocks,
           "ensemble_cell_steps": work_e, "sequential_cell_steps": work_s,
           "ensemble_cell_steps_per_s": work_e / el_e,
           "sequential_cell_steps_per_s": work_s / el_s,
           "speedup": (work_e / el_e) / (work_s / el_s)}
    log(str(out))
    return out
