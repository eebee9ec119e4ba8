"""The fast wave-speed path (characteristic polynomial, Sturm test, Laguerre) against the full
QR solve, on the host build of the device code."""
import numpy as np
import pytest

import cellhost as CH
from _common import gaussian_raw, mixture_raw, to35

PLANES = [(0, 1), (0, 2), (1, 0), (1, 2), (2, 0)]


def _gauss(rho, u, C):
    return to35(gaussian_raw(rho, u, np.asarray(C, float), 4))


@pytest.mark.parametrize("u,C", [
    ((0.0, 0.0, 0.0), np.eye(3)),
    ((0.1, 0.2, 0.3), np.eye(3) * 1.5),
    ((1.06, -1.06, 0.0), np.diag([1.0, 2.0, 0.5])),
    ((0.3, -0.2, 0.1), [[1.2, 0.3, 0.2], [0.3, 1.5, 0.4], [0.2, 0.4, 1.8]]),
])
def test_gaussian_planes_take_the_fast_path(u, C):
    """Gaussian planes have a double root (at the mean); the fast path must still prove
    hyperbolicity and return the QR extremes, u +- sqrt(3) sigma for uncorrelated planes."""
    M = _gauss(0.7, u, C)
    for f, g in PLANES:
        r = CH.plane_compare(M, f, g)
        assert not r["qr_cplx"] and r["fast_real"], (f, g, r)
        s = max(1.0, abs(r["qr"][0]), abs(r["qr"][1]))
        assert abs(r["fast"][0] - r["qr"][0]) <= 1e-12 * s, r
        assert abs(r["fast"][1] - r["qr"][1]) <= 1e-12 * s, r
    if np.allclose(C, np.diag(np.diag(C))):
        r = CH.plane_compare(M, 0, 1)
        sig = np.sqrt(C[0][0])
        assert abs(r["qr"][1] - (u[0] + np.sqrt(3) * sig)) < 1e-12 * 4


def test_mixture_planes_agree_with_qr():
    rng = np.random.default_rng(5)
    n_real = n_agree = n = 0
    worst = 0.0
    for _ in range(300):
        M = to35(mixture_raw(rng, spread=1.5, lam_lo=0.01))
        for f, g in PLANES:
            r = CH.plane_compare(M, f, g)
            n += 1
            if r["fast_real"]:
                n_real += 1
                assert not r["qr_cplx"], r  # never claims real where QR finds complex
                s = max(1.0, abs(r["qr"][0]), abs(r["qr"][1]))
                worst = max(worst, abs(r["fast"][0] - r["qr"][0]) / s, abs(r["fast"][1] - r["qr"][1]) / s)
            n_agree += r["fast_real"] == (not r["qr_cplx"])
    print("fast path on %d/%d planes, agreement %d, worst %.2e" % (n_real, n, n_agree, worst))
    assert worst < 1e-11
    assert n_agree >= 0.97 * n and n_real >= 0.5 * n


def test_exact_warm_start_saves_iterations():
    rng = np.random.default_rng(9)
    saved = []
    for _ in range(50):
        M = to35(mixture_raw(rng, spread=1.0, lam_lo=0.05))
        for f, g in PLANES:
            cold = CH.plane_compare(M, f, g)
            if not cold["fast_real"]:
                continue
            # standardized extremes as the warm start
            lo, hi = CH.plane_std_extremes(M, f, g)
            warm = CH.plane_compare(M, f, g, lo, hi)
            assert warm["fast_real"]
            assert abs(warm["fast"][1] - cold["fast"][1]) <= 1e-12 * max(1, abs(cold["fast"][1]))
            saved.append(cold["iters"] - warm["iters"])
    assert np.mean(saved) > 2, np.mean(saved)


def test_register_qr_equals_array_qr():
    """hqr6_reg (the QR iteration with the 6x6 matrix in registers) against hqr_core on an
    array: the same operations in the same order, so the same bits, on plane Jacobians of
    mixtures and Gaussians (real, complex and double eigenvalues), on random matrices and on
    matrices that are already triangular or block-triangular (immediate deflations)."""
    import ctypes
    lib = CH.lib()
    lib.cell_qr_both.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    rng = np.random.default_rng(11)
    mats = []
    for _ in range(150):
        M = to35(mixture_raw(rng, spread=1.5, lam_lo=0.01))
        for f, g in PLANES:
            mats.append(CH.jacobian_std(M, f, g))
    mats.append(CH.jacobian_std(_gauss(1.0, (0.1, 0.2, 0.3), np.eye(3)), 0, 1))
    for _ in range(200):
        mats.append(rng.standard_normal((6, 6)))
    T = np.triu(rng.standard_normal((6, 6)))
    mats += [T, np.diag(rng.standard_normal(6)), np.zeros((6, 6))]
    B = np.triu(rng.standard_normal((6, 6)))
    B[3, 2] = 0.7
    B[5, 4] = -1.3
    mats.append(B)
    n_cplx = 0
    for J in mats:
        J = np.ascontiguousarray(J, float)
        out = np.zeros(8)
        lib.cell_qr_both(J.ctypes.data, out.ctypes.data)
        assert out[:4].tobytes() == out[4:].tobytes(), (J, out)
        n_cplx += out[2] > 0
    assert n_cplx > 20
