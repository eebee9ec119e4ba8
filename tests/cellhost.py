"""ctypes binding of tests/native/_build/libcellhost.so (device cell numerics built for the
host; CPU tests only)."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_L = None
D = ctypes.POINTER(ctypes.c_double)


def lib():
    global _L
    if _L is None:
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "native")], check=True)
        _L = ctypes.CDLL(os.path.join(HERE, "native", "_build", "libcellhost.so"))
    return _L


def _p(a):
    return a.ctypes.data_as(D)


def _f(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def flux_closure35(M, flag2D=0, Ma=0.0):
    M = _f(M); Fx, Fy, Fz, Mr = (np.zeros(35) for _ in range(4)); fl = ctypes.c_uint32()
    lib().cell_flux_closure35(_p(M), flag2D, ctypes.c_double(Ma), _p(Fx), _p(Fy), _p(Fz), _p(Mr),
                              ctypes.byref(fl))
    return Fx, Fy, Fz, Mr, fl.value


def eigenvalues6(M, axis, flag2D=0, Ma=0.0):
    M = _f(M); Mr = np.zeros(35); a = ctypes.c_double(); b = ctypes.c_double(); fl = ctypes.c_uint32()
    lib().cell_eig6(_p(M), axis, flag2D, ctypes.c_double(Ma), ctypes.byref(a), ctypes.byref(b), _p(Mr),
                    ctypes.byref(fl))
    return a.value, b.value, Mr, fl.value


def pass1(M, flag2D=0, Ma=0.0, positivity=True):
    M = _f(M); Mr, Fx, Fy, Fz = (np.zeros(35) for _ in range(4)); vp = np.zeros(6); fl = ctypes.c_uint32()
    lib().cell_pass1(_p(M), flag2D, ctypes.c_double(Ma), int(positivity), _p(Mr), _p(Fx), _p(Fy), _p(Fz),
                     _p(vp), ctypes.byref(fl))
    return Mr, Fx, Fy, Fz, vp, fl.value


def pass2(M, dt, Kn, flag2D=0, Ma=0.0, positivity=True):
    M = _f(M); Mo = np.zeros(35); fl = ctypes.c_uint32()
    lib().cell_pass2(_p(M), flag2D, ctypes.c_double(Ma), int(positivity), ctypes.c_double(dt),
                     ctypes.c_double(Kn), _p(Mo), ctypes.byref(fl))
    return Mo, fl.value


def collision35(M, dt, Kn, positivity=True):
    M = _f(M); Mo = np.zeros(35)
    lib().cell_collision(_p(M), ctypes.c_double(dt), ctypes.c_double(Kn), int(positivity), _p(Mo))
    return Mo


def closure1d(m5):
    m = _f(m5); a = ctypes.c_double(); b = ctypes.c_double()
    lib().cell_closure1d(_p(m), ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def plane_compare(M, f, g, wmin=float("nan"), wmax=float("nan")):
    """QR and fast (Sturm + Laguerre) solves of plane (f, g): dict of both results."""
    M = _f(M); out = np.zeros(10)
    lib().cell_plane_compare(_p(M), f, g, ctypes.c_double(wmin), ctypes.c_double(wmax), _p(out))
    return {"qr": (out[0], out[1]), "qr_cplx": bool(out[2]), "fast": (out[3], out[4]),
            "fast_real": bool(out[5]), "iters": int(out[6]), "shift": out[7], "scale": out[8],
            "fast_cplx": bool(out[9])}


def jacobian_std(M, f, g):
    """Standardized 6x6 Jacobian of plane (f, g)."""
    M = _f(M); J = np.zeros(36)
    lib().cell_jacobian_std(_p(M), f, g, _p(J))
    return J.reshape(6, 6)


def fast_eigs_J(J, sig=1.0, need_vals=True):
    """The fast standardized solve on a 6x6 matrix: dict (real, cplx, lmin, lmax, iters)."""
    J = _f(J).ravel(); out = np.zeros(5)
    lib().cell_fast_eigs_J(_p(J), ctypes.c_double(sig), int(need_vals), _p(out))
    return {"real": bool(out[0]), "cplx": bool(out[1]), "lmin": out[2], "lmax": out[3],
            "iters": int(out[4])}


def plane_std_extremes(M, f, g):
    """Standardized QR extremes of plane (f, g)."""
    r = plane_compare(M, f, g)
    return tuple((v - r["shift"]) / r["scale"] for v in r["qr"])


def host_step(M, Kn, Ma, CFL=0.7, dtmax=1e-2, flag2D=0, t=0.0, tmax=1e9):
    """One device step emulated on the host: M Fortran (Nx,Ny,Nz,35) -> (M', dt)."""
    Nx, Ny, Nz = M.shape[:3]
    N = Nx * Ny * Nz
    soa = np.array(np.asarray(M, np.float64).ravel(order="F").reshape(35, N))  # a copy
    L = lib()
    L.host_step.restype = ctypes.c_double
    dt = L.host_step(_p(soa), Nx, Ny, Nz, ctypes.c_double(Kn), ctypes.c_double(Ma), flag2D,
                     ctypes.c_double(CFL), ctypes.c_double(dtmax), ctypes.c_double(1.0 / Nx),
                     ctypes.c_double(1.0 / Ny), ctypes.c_double(1.0 / Nz), ctypes.c_double(tmax),
                     ctypes.c_double(t), None)
    return soa.reshape(-1, order="C").reshape((35, N)).T.reshape((Nx, Ny, Nz, 35), order="F"), dt
