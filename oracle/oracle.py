"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, always as the checker.  Build with `make -C oracle`.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

D = ctypes.POINTER(ctypes.c_double)
U32 = ctypes.POINTER(ctypes.c_uint32)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "_build", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.orc_delta2starchol_L3.restype = ctypes.c_double
        L.orc_run.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(D)


def _f(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def m2cs4(M):
    M = _f(M); C = np.zeros(35); S = np.zeros(35)
    lib().orc_m2cs4(_p(M), _p(C), _p(S))
    return C, S


def flux_closure35(M, flag2D=0, Ma=0.0):
    M = _f(M)
    Fx, Fy, Fz, Mr = (np.zeros(35) for _ in range(4))
    fl = ctypes.c_uint32(0)
    lib().orc_flux_closure35(_p(M), ctypes.c_int(flag2D), ctypes.c_double(Ma), _p(Fx), _p(Fy),
                             _p(Fz), _p(Mr), ctypes.byref(fl))
    return Fx, Fy, Fz, Mr, fl.value


def delta2star3D(s28):
    s = _f(s28); E = np.zeros(36)
    lib().orc_delta2star3D(_p(s), _p(E))
    return E.reshape(6, 6)


def lower_bound_S220(s28):
    s = _f(s28); o = np.zeros(3)
    lib().orc_lower_bound_S220(_p(s), _p(o))
    return o


def bound_minor1(s28):
    s = _f(s28); o = np.zeros(12)
    lib().orc_bound_minor1(_p(s), _p(o))
    return o


def delta2starchol_L3(a10):
    a = _f(a10)
    return lib().orc_delta2starchol_L3(_p(a))


def rootsR(a10):
    a = _f(a10); re = np.zeros(3); im = np.zeros(3)
    lib().orc_rootsR(_p(a), _p(re), _p(im))
    return re + 1j * im


def rootsR_X_Y(a9):
    a = _f(a9); re = np.zeros(3); im = np.zeros(3)
    lib().orc_rootsR_X_Y(_p(a), _p(re), _p(im))
    return re + 1j * im


def jacobian6(m15):
    m = _f(m15); J = np.zeros(36)
    lib().orc_jacobian6(_p(m), _p(J))
    return J.reshape(6, 6)


def eigvals(A):
    A = _f(A); n = A.shape[0]
    wr = np.zeros(n); wi = np.zeros(n)
    lib().orc_eigvals(_p(A), ctypes.c_int(n), _p(wr), _p(wi))
    return wr + 1j * wi


def realizability_S2(S110, S101, S011):
    i = _f([S110, S101, S011]); o = np.zeros(4)
    lib().orc_realizability_S2(_p(i), _p(o))
    return o


def check2D(v10):
    i = _f(v10); o = np.zeros(10)
    lib().orc_check2D(_p(i), _p(o))
    return o


def realizable_3D(s28):
    i = _f(s28); o = np.zeros(28); f = ctypes.c_int(0)
    lib().orc_realizable_3D(_p(i), _p(o), ctypes.byref(f))
    return o, f.value


def eigenvalues6(M, axis, flag2D=0, Ma=0.0):
    M = _f(M); Mr = np.zeros(35)
    vmin = ctypes.c_double(); vmax = ctypes.c_double(); fl = ctypes.c_uint32()
    lib().orc_eigenvalues6(_p(M), ctypes.c_int(axis), ctypes.c_int(flag2D), ctypes.c_double(Ma),
                           ctypes.byref(vmin), ctypes.byref(vmax), _p(Mr), ctypes.byref(fl))
    return vmin.value, vmax.value, Mr, fl.value


def eigenvalues6z(M, flag2D=0, Ma=0.0, positivity=True):
    M = _f(M); Mr = np.zeros(35)
    vmin = ctypes.c_double(); vmax = ctypes.c_double(); fl = ctypes.c_uint32()
    lib().orc_eigenvalues6z(_p(M), ctypes.c_int(flag2D), ctypes.c_double(Ma),
                            ctypes.c_int(1 if positivity else 0), ctypes.byref(vmin),
                            ctypes.byref(vmax), _p(Mr), ctypes.byref(fl))
    return vmin.value, vmax.value, Mr, fl.value


def realizable_3D_flags(s28):
    i = _f(s28); o = np.zeros(28); f = ctypes.c_int(0); fl = ctypes.c_uint32(0)
    lib().orc_realizable_3D_flags(_p(i), _p(o), ctypes.byref(f), ctypes.byref(fl))
    return o, f.value, fl.value


def moments5_3d(M):
    """Moments5_3D: (M5, C5, S5), 21 each, in the reference's order."""
    M = _f(M); a, b, c = np.zeros(21), np.zeros(21), np.zeros(21)
    lib().orc_moments5_3d(_p(M), _p(a), _p(b), _p(c))
    return a, b, c


def closure_and_eigenvalues(mom5):
    m = _f(mom5); o = np.zeros(3)
    lib().orc_closure_and_eigenvalues(_p(m), _p(o))
    return o[0], o[1], o[2]


def initialize_m4_35(rho, u, v, w, C200, C110, C101, C020, C011, C002):
    i = _f([rho, u, v, w, C200, C110, C101, C020, C011, C002]); o = np.zeros(35)
    lib().orc_initialize_m4_35(_p(i), _p(o))
    return o


def s4toc4(C2, S35):
    C2 = _f(C2); S = _f(S35); o = np.zeros(35)
    lib().orc_s4toc4(_p(C2), _p(S), _p(o))
    return o


def collision35(M, dt, Kn, positivity=True):
    M = _f(M); o = np.zeros(35)
    lib().orc_collision35(_p(M), ctypes.c_double(dt), ctypes.c_double(Kn),
                          ctypes.c_int(1 if positivity else 0), _p(o))
    return o


def pas_hll(M, F, dt, dx, vpmin, vpmax):
    M = _f(M); F = _f(F); a = _f(vpmin); b = _f(vpmax)
    Np = M.shape[0]; o = np.zeros_like(M)
    lib().orc_pas_hll(_p(M), _p(F), ctypes.c_int(Np), ctypes.c_double(dt), ctypes.c_double(dx),
                      _p(a), _p(b), _p(o))
    return o


def pass1_batch(M, flag2D=0, Ma=0.0, positivity=True, nthreads=None):
    """M: [n][35].  Returns Mr, Fx, Fy, Fz ([n][35]), vp [n][6], flags [n]."""
    M = _f(M); n = M.shape[0]
    Mr, Fx, Fy, Fz = (np.zeros((n, 35)) for _ in range(4))
    vp = np.zeros((n, 6)); fl = np.zeros(n, dtype=np.uint32)
    nt = nthreads or os.cpu_count()
    lib().orc_pass1_batch(_p(M), ctypes.c_longlong(n), ctypes.c_int(flag2D), ctypes.c_double(Ma),
                          ctypes.c_int(1 if positivity else 0), _p(Mr), _p(Fx), _p(Fy), _p(Fz),
                          _p(vp), fl.ctypes.data_as(U32), ctypes.c_int(nt))
    return Mr, Fx, Fy, Fz, vp, fl


def run(M, Nx, Ny, Nz, Kn, Ma, flag2D, CFL, dtmax, dx, dy, dz, positivity, tmax, nnmax,
        t0=0.0, nn0=0, nthreads=None):
    """Step loop on a Julia-layout state (Nx,Ny,Nz,35) given as a Fortran-ordered array.
    Returns (M_final, t, nn, dts, rc)."""
    Mf = np.asfortranarray(np.array(M, dtype=np.float64))
    buf = Mf.ravel(order="F").copy()
    t = ctypes.c_double(t0); nn = ctypes.c_longlong(nn0)
    dts = np.zeros(max(int(nnmax) - nn0, 1))
    nt = nthreads or os.cpu_count()
    rc = lib().orc_run(_p(buf), Nx, Ny, Nz, ctypes.c_double(Kn), ctypes.c_double(Ma), flag2D,
                       ctypes.c_double(CFL), ctypes.c_double(dtmax), ctypes.c_double(dx),
                       ctypes.c_double(dy), ctypes.c_double(dz), ctypes.c_int(1 if positivity else 0),
                       ctypes.c_double(tmax), ctypes.c_longlong(int(nnmax)), ctypes.byref(t),
                       ctypes.byref(nn), ctypes.c_int(nt), _p(dts))
    out = buf.reshape((Nx, Ny, Nz, 35), order="F")
    return out, t.value, nn.value, dts[: nn.value - nn0], rc
