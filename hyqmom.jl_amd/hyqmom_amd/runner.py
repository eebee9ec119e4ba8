"""simulation_runner(params) and run_simulation_with_snapshots(params; snapshot_interval):
the reference's Julia API for the 3D step (src/simulation_runner.jl:63-759, 994-1011),
mirrored over the C ABI of libhyqmom_hip.so.

params is a dict (or any mapping / object with attributes) with the reference's field
names: Nx, Ny, Nz, tmax, Kn, Ma, flag2D, CFL, Nmom, nnmax, dtmax, rhol, rhor, T, r110,
r101, r011, symmetry_check_interval, homogeneous_z, debug_output; optional positivity,
snapshot_interval, snapshot_filename, xmin..zmax, use_custom_ic, ic_background, ic_jets,
and (this build) precision in {"fp64", "mixed", "mixed_newton", "mixed_learned"} with
warmstart_model (a warmstart.WarmStartMLP or a path; default models/warmstart_mlp.f32 when
present) for "mixed_learned".
"""
import ctypes
import math
import os
import time

import numpy as np

from . import _lib
from .initial_conditions import cell_centers, initialize_moment_field
from .moments import maxwellian35

REQUIRED = ["Nx", "Ny", "Nz", "tmax", "Kn", "Ma", "flag2D", "CFL", "Nmom", "nnmax", "dtmax",
            "rhol", "rhor", "T", "r110", "r101", "r011"]


def _get(params, key, default=None):
    if isinstance(params, dict):
        return params.get(key, default)
    return getattr(params, key, default)


def hardcoded_jets(Nx, Ny, Nz, rhol, rhor, T, Ma, r110=0.0, r101=0.0, r011=0.0):
    """The original crossing-jets IC (src/simulation_runner.jl:200-271)."""
    C = np.array([[T, r110 * math.sqrt(T * T), r101 * math.sqrt(T * T)],
                  [r110 * math.sqrt(T * T), T, r011 * math.sqrt(T * T)],
                  [r101 * math.sqrt(T * T), r011 * math.sqrt(T * T), T]])
    Uc = Ma / math.sqrt(2.0)
    bg = maxwellian35(rhor, (0.0, 0.0, 0.0), C)
    Mt = maxwellian35(rhol, (-Uc, -Uc, 0.0), C)
    Mb = maxwellian35(rhol, (Uc, Uc, 0.0), C)
    Cx, Cy = int(math.floor(0.1 * Nx)), int(math.floor(0.1 * Ny))
    Mint_x, Maxt_x, Minb_x, Maxb_x = Nx // 2 + 1, Nx // 2 + 1 + Cx, Nx // 2 - Cx, Nx // 2
    Mint_y, Maxt_y, Minb_y, Maxb_y = Ny // 2 + 1, Ny // 2 + 1 + Cy, Ny // 2 - Cy, Ny // 2
    Cz = min(Cx, Cy)
    Mint_z, Maxt_z = Nz // 2 + 1 - Cz // 2, Nz // 2 + 1 + Cz // 2
    gi = np.arange(1, Nx + 1)[:, None, None]
    gj = np.arange(1, Ny + 1)[None, :, None]
    gk = np.arange(1, Nz + 1)[None, None, :]
    M = np.empty((Nx, Ny, Nz, 35), order="F")
    M[...] = bg
    zin = (gk >= Mint_z) & (gk <= Maxt_z)
    bot = (gi >= Minb_x) & (gi <= Maxb_x) & (gj >= Minb_y) & (gj <= Maxb_y) & zin
    top = (gi >= Mint_x) & (gi <= Maxt_x) & (gj >= Mint_y) & (gj <= Maxt_y) & zin
    M[bot] = Mb
    M[top] = Mt
    return M


def initial_state(params):
    Nx, Ny, Nz = int(_get(params, "Nx")), int(_get(params, "Ny")), int(_get(params, "Nz"))
    r = [float(_get(params, k, 0.0)) for k in ("r110", "r101", "r011")]
    if _get(params, "use_custom_ic", False):
        M, _ = initialize_moment_field(Nx, Ny, Nz, _get(params, "ic_background"), _get(params, "ic_jets"),
                                       *r, xmin=_get(params, "xmin", -0.5), xmax=_get(params, "xmax", 0.5),
                                       ymin=_get(params, "ymin", -0.5), ymax=_get(params, "ymax", 0.5),
                                       zmin=_get(params, "zmin", -0.5), zmax=_get(params, "zmax", 0.5))
        return M
    return hardcoded_jets(Nx, Ny, Nz, float(_get(params, "rhol")), float(_get(params, "rhor")),
                          float(_get(params, "T")), float(_get(params, "Ma")), *r)


class Solver:
    """One device context: state, step and run on the GPU."""

    def __init__(self, params):
        for k in REQUIRED:
            if _get(params, k) is None:
                raise ValueError("params is missing required field %s" % k)
        if int(_get(params, "Nmom")) != 35:
            raise ValueError("Nmom must be 35")
        self.params = params
        p = _lib.HyqParams()
        p.Nx, p.Ny, p.Nz = int(_get(params, "Nx")), int(_get(params, "Ny")), int(_get(params, "Nz"))
        p.Nmom = 35
        p.flag2D = int(_get(params, "flag2D"))
        p.positivity = 1 if _get(params, "positivity", True) else 0
        p.precision = _lib.PREC[_get(params, "precision", "fp64")]
        p.tmax = float(_get(params, "tmax"))
        p.Kn = float(_get(params, "Kn"))
        p.Ma = float(_get(params, "Ma"))
        p.CFL = float(_get(params, "CFL"))
        p.dtmax = float(_get(params, "dtmax"))
        p.xmin, p.xmax = float(_get(params, "xmin", -0.5)), float(_get(params, "xmax", 0.5))
        p.ymin, p.ymax = float(_get(params, "ymin", -0.5)), float(_get(params, "ymax", 0.5))
        p.zmin, p.zmax = float(_get(params, "zmin", -0.5)), float(_get(params, "zmax", 0.5))
        self.p = p
        self.shape = (p.Nx, p.Ny, p.Nz, 35)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().hyq_create(ctypes.byref(p), ctypes.byref(h)), "hyq_create")
        self.h = h
        self.t = 0.0
        self.nn = 0
        self._mlp = None
        if p.precision == _lib.PREC["mixed_learned"]:
            model = _get(params, "warmstart_model")
            if model is None or isinstance(model, str):
                from .warmstart import DEFAULT_MODEL, WarmStartMLP
                path = model or DEFAULT_MODEL
                model = WarmStartMLP.load(path) if os.path.exists(path) else None
            if model is not None:
                self.set_warm_start(model)

    def set_warm_start(self, model):
        """Attach a warmstart.WarmStartMLP (precision "mixed_learned"); None detaches and the
        root iteration then starts from the analytic root bounds."""
        _lib.check(_lib.lib().hyq_set_mlp(self.h, model.handle if model is not None else None),
                   "hyq_set_mlp")
        self._mlp = model

    def close(self):
        if self.h:
            _lib.lib().hyq_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, M):
        buf = np.asfortranarray(np.asarray(M, dtype=np.float64).reshape(self.shape, order="F"))
        flat = buf.ravel(order="F")
        _lib.check(_lib.lib().hyq_set_state(self.h, flat.ctypes.data_as(_lib.D)), "hyq_set_state")

    def get_state(self):
        flat = np.empty(int(np.prod(self.shape)), dtype=np.float64)
        _lib.check(_lib.lib().hyq_get_state(self.h, flat.ctypes.data_as(_lib.D)), "hyq_get_state")
        return flat.reshape(self.shape, order="F")

    def step(self):
        t = ctypes.c_double(self.t); dt = ctypes.c_double(); nn = _lib.I64(self.nn)
        rc = _lib.lib().hyq_step(self.h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(nn))
        if rc in (0, -4):  # HYQ_ENONFINITE still completes the device step
            self.t, self.nn = t.value, nn.value
        _lib.check(rc, "hyq_step")
        return dt.value

    def run(self, tmax, nnmax):
        t = ctypes.c_double(self.t); nn = _lib.I64(self.nn)
        rc = _lib.lib().hyq_run(self.h, ctypes.c_double(tmax), _lib.I64(int(nnmax)), ctypes.byref(t),
                                ctypes.byref(nn))
        if rc in (0, -4):  # HYQ_ENONFINITE still completes the device steps
            self.t, self.nn = t.value, nn.value
        _lib.check(rc, "hyq_run")

    def wave_speeds(self):
        n = self.p.Nx * self.p.Ny * self.p.Nz
        out = np.empty(6 * n)
        _lib.check(_lib.lib().hyq_get_wave_speeds(self.h, out.ctypes.data_as(_lib.D)), "wave speeds")
        return out.reshape(6, n)

    def flags(self):
        n = self.p.Nx * self.p.Ny * self.p.Nz
        out = np.empty(n, dtype=np.uint32)
        _lib.check(_lib.lib().hyq_get_flags(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
                   "flags")
        return out

    def kernel_times(self):
        """(pass 1 ms, update ms, pass 2 ms, steps) averaged since the last reset."""
        a, b, d = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        n = _lib.I64()
        _lib.check(_lib.lib().hyq_kernel_times(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(d),
                                               ctypes.byref(n)), "kernel times")
        return a.value, b.value, d.value, n.value

    def reset_kernel_times(self):
        _lib.check(_lib.lib().hyq_reset_kernel_times(self.h), "reset times")

    def deferred_cells(self):
        """(pass 1, pass 2) cells of the last step that the fast modes sent to the full path."""
        a, b = _lib.I64(), _lib.I64()
        _lib.check(_lib.lib().hyq_deferred_cells(self.h, ctypes.byref(a), ctypes.byref(b)), "deferred cells")
        return a.value, b.value

    def fast_path_counts(self):
        """Why cells of the last step left the fast kernels (fast modes): a dict with, per
        pass, the cells with plane solves, the plane solves, the cells a plane solve found
        complex (eigenvalue correction, full path) and the cells on the full path from the
        start (z density clamp or re-projection, non-finite plane)."""
        a = (_lib.I64 * 8)()
        _lib.check(_lib.lib().hyq_fast_path_counts(self.h, a), "fast path counts")
        keys = ("plane_solve_cells", "plane_solves", "complex_cells", "full_path_cells")
        return {"pass1": dict(zip(keys, a[:4])), "pass2": dict(zip(keys, a[4:]))}

    def snapshot_fields(self):
        """(C, S): the central and standardized moment fields of the current state
        (compute_central_field / compute_standardized_field), computed on the device, in the
        state's (Nx, Ny, Nz, 35) layout."""
        C = np.empty(int(np.prod(self.shape)), dtype=np.float64)
        S = np.empty_like(C)
        _lib.check(_lib.lib().hyq_snapshot_fields(self.h, C.ctypes.data_as(_lib.D), S.ctypes.data_as(_lib.D)),
                   "snapshot fields")
        return C.reshape(self.shape, order="F"), S.reshape(self.shape, order="F")

    def symmetry_maxdiff(self):
        """(MaxDiff, per-moment values) of the runner's symmetry check
        (simulation_runner.jl:565-628) on the current state, computed on the device."""
        m = ctypes.c_double()
        per = np.zeros(5)
        _lib.check(_lib.lib().hyq_symmetry_maxdiff(self.h, ctypes.byref(m), per.ctypes.data_as(_lib.D)),
                   "symmetry")
        return m.value, per

    def vmax_breakdown(self):
        """((vx_max, vy_max, vz_max), (i, j, k) 0-based of the fastest cell) of the last step's
        wave speeds, as the reference's first-step diagnostics (simulation_runner.jl:404-494)."""
        v = np.zeros(3)
        w = (_lib.I64 * 3)()
        _lib.check(_lib.lib().hyq_vmax_breakdown(self.h, v.ctypes.data_as(_lib.D), w), "vmax breakdown")
        return tuple(v), (w[0], w[1], w[2])


def grid(params):
    """grid_out of simulation_runner.jl:720-730."""
    Nx, Ny, Nz = int(_get(params, "Nx")), int(_get(params, "Ny")), int(_get(params, "Nz"))
    lo = [float(_get(params, k, d)) for k, d in (("xmin", -.5), ("ymin", -.5), ("zmin", -.5))]
    hi = [float(_get(params, k, d)) for k, d in (("xmax", .5), ("ymax", .5), ("zmax", .5))]
    out = {}
    for ax, n, a, b in zip("xyz", (Nx, Ny, Nz), lo, hi):
        out[ax] = np.linspace(a, b, n + 1)
        out["d" + ax] = (b - a) / n
        out[ax + "m"] = cell_centers(a, b, n)[0]
    return out


def simulation_runner(params, verbose=False):
    """Returns (M_final (Nx,Ny,Nz,35), final_time, time_steps, grid_out) as the reference
    (src/simulation_runner.jl:63-759), or (snapshot_filename, grid_out) in snapshot mode."""
    snap = int(_get(params, "snapshot_interval", 0) or 0)
    solver = Solver(params)
    try:
        solver.set_state(initial_state(params))
        tmax = float(_get(params, "tmax"))
        nnmax = int(_get(params, "nnmax"))
        if snap > 0:
            from .snapshots import SnapshotWriter
            fname = _get(params, "snapshot_filename") or "snapshots.npz"
            w = SnapshotWriter(fname, params, snap)
            w.write(solver.get_state(), solver.t, solver.nn, *solver.snapshot_fields())
            while solver.t < tmax and solver.nn < nnmax:
                solver.run(tmax, min(nnmax, solver.nn + snap))
                w.write(solver.get_state(), solver.t, solver.nn, *solver.snapshot_fields())
            w.close(grid(params))
            return fname, grid(params)
        t0 = time.time()
        sym = int(_get(params, "symmetry_check_interval", 0) or 0)
        while solver.t < tmax and solver.nn < nnmax:
            t_prev = solver.t
            solver.step()
            if verbose:
                n = params_cells(params)
                el = time.time() - t0
                if solver.nn == 1:
                    _print_first_step(solver, params)
                line = "Step %4d: t = %.4f, dt = %.4e, wall = %.4f s, s/pt = %.4e s" % (
                    solver.nn, solver.t, solver.t - t_prev, el, el / n)
                if sym > 0 and (solver.nn % sym == 0 or solver.nn == 1):
                    line += ", MaxDiff = %.3e" % solver.symmetry_maxdiff()[0]
                print(line)
                t0 = time.time()
        return solver.get_state(), solver.t, solver.nn, grid(params)
    finally:
        solver.close()


def _print_first_step(solver, params):
    """The vmax breakdown the reference prints after its first step (simulation_runner.jl:
    404-427), from the device reduction."""
    (vx, vy, vz), (i, j, k) = solver.vmax_breakdown()
    vmax = max(vx, vy, vz)
    print("  first step: vx_max = %.6e, vy_max = %.6e, vz_max = %.6e, vmax = %.6e, "
          "fastest cell (%d,%d,%d)" % (vx, vy, vz, vmax, i + 1, j + 1, k + 1))


def params_cells(params):
    return int(_get(params, "Nx")) * int(_get(params, "Ny")) * int(_get(params, "Nz"))


def run_simulation(Np=None, Nx=20, Ny=20, Nz=1, tmax=0.1, num_workers=1, verbose=True,
                   save_output=False, enable_plots=False, save_figures=False, output_dir=".",
                   Kn=1.0, Ma=0.0, flag2D=0, CFL=0.5, homogeneous_z=True, debug_output=False,
                   precision="fp64"):
    """src/simulation_runner.jl:816-944: the original crossing jets (rhol = 1, rhor = 0.01,
    T = 1, no correlations) on the unit cube, dtmax = CFL min(dx, dy, dz), run on the device.
    Returns the reference's dictionary (M, final_time, time_steps, xm, ym, zm, Nx, Ny, Nz,
    tmax; string keys for its symbols).  num_workers, plots and output files have no role on
    one device; they are accepted and ignored."""
    if Np is not None:
        Nx = Ny = Np
    dx, dy, dz = 1.0 / Nx, 1.0 / Ny, 1.0 / Nz
    dtmax = CFL * min(dx, dy, dz)
    params = dict(Nx=Nx, Ny=Ny, Nz=Nz, tmax=tmax, Kn=Kn, Ma=Ma, flag2D=flag2D, CFL=CFL, dx=dx,
                  dy=dy, dz=dz, Nmom=35, nnmax=math.ceil(tmax / dtmax) + 100000, dtmax=dtmax,
                  rhol=1.0, rhor=0.01, T=1.0, r110=0.0, r101=0.0, r011=0.0,
                  symmetry_check_interval=10, homogeneous_z=homogeneous_z,
                  enable_memory_tracking=False, debug_output=debug_output, precision=precision)
    if verbose:
        print("HyQMOM simulation: grid %dx%dx%d, tmax %g, Kn %g, Ma %g, CFL %g, dtmax %g"
              % (Nx, Ny, Nz, tmax, Kn, Ma, CFL, dtmax))
    M, t, nn, g = simulation_runner(params, verbose=verbose)
    if verbose:
        print("Simulation complete: final time %g, %d time steps" % (t, nn))
    return {"M": M, "final_time": t, "time_steps": nn, "xm": g["xm"], "ym": g["ym"],
            "zm": g["zm"], "Nx": Nx, "Ny": Ny, "Nz": Nz, "tmax": tmax}


def run_simulation_with_snapshots(params, snapshot_interval=2):
    """src/simulation_runner.jl:994-1011"""
    p = dict(params) if isinstance(params, dict) else dict(vars(params))
    p["snapshot_interval"] = snapshot_interval
    return simulation_runner(p)
