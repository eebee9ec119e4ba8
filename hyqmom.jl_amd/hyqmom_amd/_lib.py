"""ctypes binding of libhyqmom_hip.so (include/hyqmom.h).

The library is built in-tree by hyqmom.jl_amd/build.py (hipcc, gfx950).  Loading fails
loudly when it is missing: there is no CPU fallback for the device step.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("HYQMOM_AMD_LIB") or os.path.join(os.path.dirname(HERE), "build", "libhyqmom_hip.so")

HYQ_OK = 0
HYQ_ENONFINITE = -4
ERRORS = {-1: "HYQ_EINVAL", -2: "HYQ_ENOMEM", -3: "HYQ_EHIP", -4: "HYQ_ENONFINITE",
          -5: "HYQ_ENOTCONV", -6: "HYQ_ENODEV"}
PREC = {"fp64": 0, "mixed": 1, "mixed_learned": 2, "mixed_newton": 3}

FLAG_NAMES = ["UNIVARIATE", "EDGE_CORNER", "FACES", "FLAG220", "ROOTS2D", "ROOTS4", "EIGCORR_X",
              "EIGCORR_Y", "EIGCORR_Z", "ZNEGVAR", "NONFINITE", "RHOCLAMP", "EIG_FALLBACK",
              "QR_NOCONV"]


def newton_iterations(flags):
    """Laguerre iterations per cell (flag bits 16..27) of the fast wave-speed modes."""
    return (flags >> 16) & 0xFFF


class HyqParams(ctypes.Structure):
    _fields_ = [("Nx", ctypes.c_int32), ("Ny", ctypes.c_int32), ("Nz", ctypes.c_int32),
                ("Nmom", ctypes.c_int32), ("flag2D", ctypes.c_int32), ("positivity", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("tmax", ctypes.c_double), ("Kn", ctypes.c_double), ("Ma", ctypes.c_double),
                ("CFL", ctypes.c_double), ("dtmax", ctypes.c_double),
                ("xmin", ctypes.c_double), ("xmax", ctypes.c_double), ("ymin", ctypes.c_double),
                ("ymax", ctypes.c_double), ("zmin", ctypes.c_double), ("zmax", ctypes.c_double)]


class HyqError(RuntimeError):
    pass


_LIB = None
D = ctypes.POINTER(ctypes.c_double)
VP = ctypes.c_void_p
I64 = ctypes.c_int64

# exported symbol -> (argtypes, restype)
SIGNATURES = {
    "hyq_create": ([ctypes.POINTER(HyqParams), ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "hyq_destroy": ([VP], ctypes.c_int),
    "hyq_set_state": ([VP, D], ctypes.c_int),
    "hyq_get_state": ([VP, D], ctypes.c_int),
    "hyq_step": ([VP, D, D, ctypes.POINTER(I64)], ctypes.c_int),
    "hyq_run": ([VP, ctypes.c_double, I64, D, ctypes.POINTER(I64)], ctypes.c_int),
    "hyq_get_wave_speeds": ([VP, D], ctypes.c_int),
    "hyq_get_flags": ([VP, ctypes.POINTER(ctypes.c_uint32)], ctypes.c_int),
    "hyq_kernel_times": ([VP, D, D, D, ctypes.POINTER(I64)], ctypes.c_int),
    "hyq_reset_kernel_times": ([VP], ctypes.c_int),
    "hyq_deferred_cells": ([VP, ctypes.POINTER(I64), ctypes.POINTER(I64)], ctypes.c_int),
    "hyq_fast_path_counts": ([VP, ctypes.POINTER(I64)], ctypes.c_int),
    "hyq_snapshot_fields": ([VP, D, D], ctypes.c_int),
    "hyq_symmetry_maxdiff": ([VP, D, D], ctypes.c_int),
    "hyq_vmax_breakdown": ([VP, D, ctypes.POINTER(I64)], ctypes.c_int),
    "hyq_closure_batch": ([VP, I64, ctypes.c_double, ctypes.c_int32, VP, VP, VP, VP, VP], ctypes.c_int),
    "hyq_moments5_batch": ([VP, I64, VP, VP, VP], ctypes.c_int),
    "hyq_wavespeed_batch": ([VP, I64, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                             VP, VP, VP, VP, VP, VP], ctypes.c_int),
    "hyq_wavespeed_batch_noalloc": ([VP, I64, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                     VP, VP, VP, VP, VP, VP], ctypes.c_int),
    "hyq_eig6_batch": ([VP, I64, VP, VP], ctypes.c_int),
    "hyq_collision_batch": ([VP, I64, ctypes.c_double, ctypes.c_double, ctypes.c_int32, VP], ctypes.c_int),
    "hyq_closure1d_batch": ([VP, I64, VP, VP], ctypes.c_int),
    "hyq_hll_lines": ([VP, VP, VP, VP, I64, ctypes.c_int32, ctypes.c_double, ctypes.c_double, VP], ctypes.c_int),
    "hyq_mlp_blob_size": ([], ctypes.c_int64),
    "hyq_mlp_param_count": ([], ctypes.c_int64),
    "hyq_mlp_create": ([VP, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "hyq_mlp_create_hidden": ([ctypes.c_int32, VP, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "hyq_mlp_blob_size_hidden": ([ctypes.c_int32], ctypes.c_int64),
    "hyq_mlp_param_count_hidden": ([ctypes.c_int32], ctypes.c_int64),
    "hyq_mlp_hidden": ([VP], ctypes.c_int32),
    "hyq_mlp_destroy": ([VP], ctypes.c_int),
    "hyq_mlp_get_blob": ([VP, VP], ctypes.c_int),
    "hyq_mlp_set_blob": ([VP, VP], ctypes.c_int),
    "hyq_mlp_infer": ([VP, VP, I64, VP], ctypes.c_int),
    "hyq_mlp_infer_layers": ([VP, VP, I64, VP], ctypes.c_int),
    "hyq_mlp_set_shape": ([VP, ctypes.c_int32], ctypes.c_int),
    "hyq_mlp_grad": ([VP, VP, VP, I64, VP, VP], ctypes.c_int),
    "hyq_mlp_train_step": ([VP, VP, VP, I64, ctypes.c_float, VP], ctypes.c_int),
    "hyq_mlp_last_error": ([], ctypes.c_char_p),
    "hyq_set_mlp": ([VP, VP], ctypes.c_int),
    "hyq_plane_dataset": ([VP, I64, VP, VP, VP], ctypes.c_int),
    "hyq_ensemble_create": ([ctypes.POINTER(HyqParams), ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)],
                            ctypes.c_int),
    "hyq_ensemble_destroy": ([VP], ctypes.c_int),
    "hyq_ensemble_set_state": ([VP, ctypes.c_int32, D], ctypes.c_int),
    "hyq_ensemble_get_state": ([VP, ctypes.c_int32, D], ctypes.c_int),
    "hyq_ensemble_status": ([VP, D, D, ctypes.POINTER(I64)], ctypes.c_int),
    "hyq_ensemble_run": ([VP, I64, ctypes.c_int32], ctypes.c_int),
    "hyq_ensemble_set_blocks": ([VP, ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
    "hyq_ensemble_tune": ([VP, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), D], ctypes.c_int),
    "hyq_ensemble_last_error": ([], ctypes.c_char_p),
    "hyq_last_error": ([], ctypes.c_char_p),
    "hyq_version": ([], ctypes.c_char_p),
}


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise HyqError("libhyqmom_hip.so is not built (%s): run __graft_entry__.build()" % LIBPATH)
        # torch ships its own HIP runtime under the same soname as /opt/rocm's: when both
        # are used in one process, torch's must be loaded first so that this library
        # resolves to it (the other order leaves torch without a device)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIBPATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def check(rc, what=""):
    if rc != HYQ_OK:
        msg = lib().hyq_last_error().decode()
        raise HyqError("%s failed: %s (%s)" % (what, ERRORS.get(rc, rc), msg))
