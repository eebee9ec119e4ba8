"""The C ABI library loads and exports every symbol include/hyqmom.h declares (CPU)."""
import ctypes
import os
import re

from _common import HERE

ROOT = os.path.dirname(HERE)


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "hyqmom.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|int32_t|const char\*)\s+(hyq_\w+)\s*\(", txt, re.M)))


def test_library_exports_header_symbols():
    from hyqmom_amd import _lib
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_version_and_error_without_gpu():
    from hyqmom_amd import _lib
    L = _lib.lib()
    assert b"gfx950" in L.hyq_version()
    assert L.hyq_closure_batch(None, 1, 0.0, 0, None, None, None, None, None) == -1
    assert b"bad argument" in L.hyq_last_error()


def test_invalid_params_rejected():
    from hyqmom_amd import _lib
    L = _lib.lib()
    p = _lib.HyqParams(Nx=4, Ny=4, Nz=4, Nmom=34)
    h = ctypes.c_void_p()
    assert L.hyq_create(ctypes.byref(p), ctypes.byref(h)) == -1


def test_cell_limit_rejected_before_any_allocation():
    """The update addresses a plane with 32-bit byte offsets, so a context, an ensemble (all
    its runs together) and an HLL line hold at most 2^29 - 1 cells; larger requests are
    rejected before anything touches the GPU."""
    from hyqmom_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    p = _lib.HyqParams(Nx=1024, Ny=1024, Nz=512, Nmom=35)  # 2^29 cells
    assert L.hyq_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert b"2^29" in L.hyq_last_error()
    runs = (_lib.HyqParams * 2)(_lib.HyqParams(Nx=512, Ny=512, Nz=1024, Nmom=35),
                                _lib.HyqParams(Nx=512, Ny=512, Nz=1024, Nmom=35))
    assert L.hyq_ensemble_create(runs, 2, ctypes.byref(h)) == -1
    assert b"2^29" in L.hyq_ensemble_last_error()
    d = ctypes.c_void_p(8)  # never dereferenced: the length is checked first
    assert L.hyq_hll_lines(d, d, d, d, 1, 1 << 29, 1e-3, 0.01, d) == -1
    assert b"2^29" in L.hyq_last_error()


def test_network_width_matches_cpu_reference():
    """The library's parameter blob has the layout of the CPU reference network (same hidden
    width), so blobs move between them (a host-only query, no GPU)."""
    import mlp_ref as R
    from hyqmom_amd import _lib
    from hyqmom_amd.warmstart import hidden_width
    L = _lib.lib()
    assert L.hyq_mlp_param_count() == R.NPARAM
    assert L.hyq_mlp_blob_size() == R.NBLOB
    assert hidden_width(R.NPARAM) == R.HID
    for H in (32, 64, 96, 128):
        assert hidden_width(2 * H * H + 35 * H + 16) == H
