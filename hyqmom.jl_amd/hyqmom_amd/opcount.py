"""Algorithmic FLOPs of the cell numerics on given states (libhyqmom_opcount.so: the device
cell code compiled for the host with every FP64 operation counted)."""
import ctypes
import os

import numpy as np

from ._lib import HERE, HyqError

LIBPATH = os.path.join(os.path.dirname(HERE), "build", "libhyqmom_opcount.so")
_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIBPATH):
            raise HyqError("libhyqmom_opcount.so is not built (%s)" % LIBPATH)
        L = ctypes.CDLL(LIBPATH)
        L.hyq_opcount.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.hyq_opcount.restype = ctypes.c_int
        _L = L
    return _L


def flops(cells, flag2D=0, Ma=0.0, dt=1e-3, Kn=1.0, newton=False):
    """cells: [n][35] raw moments.  Returns (pass-1 FLOPs, pass-2 FLOPs) summed over cells."""
    soa = np.ascontiguousarray(np.asarray(cells, dtype=np.float64).T)
    a, b = ctypes.c_double(), ctypes.c_double()
    rc = lib().hyq_opcount(soa.ctypes.data_as(ctypes.c_void_p), soa.shape[1], int(flag2D), float(Ma),
                           float(dt), float(Kn), 1 if newton else 0, ctypes.byref(a), ctypes.byref(b))
    if rc != 0:
        raise HyqError("hyq_opcount: bad arguments")
    return a.value, b.value


def state_flops_per_cell(M, stride=7, **kw):
    """Mean FLOPs per cell of pass 1 and pass 2 over every stride-th cell of a (Nx,Ny,Nz,35)
    state."""
    flat = np.asarray(M).reshape(-1, 35, order="F")
    sample = flat[::stride]
    a, b = flops(sample, **kw)
    return a / len(sample), b / len(sample)
