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



# free entries of the plane Jacobian (rows 0, 1 and 3 are shift rows; J[2][4], J[2][5] and
# J[5][2] are zero), and the six solved for to give a prescribed spectrum
_FREE = [(2, 0), (2, 1), (2, 2), (2, 3), (4, 0), (4, 1), (4, 2), (4, 3), (4, 4), (4, 5), (5, 0),
         (5, 1), (5, 3), (5, 4), (5, 5)]
_SOLVED = [(2, 0), (2, 1), (2, 2), (2, 3), (4, 0), (5, 0)]


def _structured_jacobian(rng, lam):
    """A matrix with the plane Jacobian's sparsity and characteristic polynomial prod(l - lam)
    (random free entries, six of them solved for), or None."""
    from scipy.optimize import fsolve
    target = np.real(np.poly(lam))[1:]
    for _ in range(20):
        J = np.zeros((6, 6))
        J[0, 1] = J[1, 2] = J[3, 4] = 1.0
        for ij in _FREE:
            J[ij] = 0.6 * rng.standard_normal()

        def resid(x):
            K = J.copy()
            for ij, v in zip(_SOLVED, x):
                K[ij] = v
            return np.real(np.poly(K))[1:] - target

        x, _, ier, _ = fsolve(resid, [J[ij] for ij in _SOLVED], full_output=True, xtol=1e-15)
        if ier == 1 and np.abs(resid(x)).max() < 1e-13:
            for ij, v in zip(_SOLVED, x):
                J[ij] = v
            return J
    return None


def test_complex_pair_decisions_follow_the_reference_test():
    """Matrices of the plane Jacobian's form with eigenvalues -1.7, -0.9, 0.35 +- i delta,
    1.1, 1.6, delta from 0 to 0.1, and with the pair split along the real axis instead.  The
    reference corrects a plane when LAPACK's eigenvalues have |imag| > 1000 eps.  The fast solve may claim the correction only where LAPACK finds
    the plane complex for the matrix and for 16 perturbations of its entries by up to 4 ulps.
    It may prove the plane real, with LAPACK's extremes, only where LAPACK finds it real, for
    the matrix and for 16 perturbations of its entries by up to 1e-13: a pair whose
    discriminant is below the characteristic polynomial's resolution (|imag| up to ~3e-8) is
    taken as a double real root, as the reference's exact Maxwellian planes are, and LAPACK's
    own decision on such a pair is set by perturbations of that size.  (A pair this close in
    a matrix of this form is nearly a Jordan block, where inverse iteration converges slowly:
    the 2x2 block test must not be misled by an unconverged subspace.)"""
    rng = np.random.default_rng(99)
    thr = 1000 * np.finfo(float).eps

    def lapack_imag(J, rel):
        return [np.abs(np.linalg.eigvals(J * (1 + rng.uniform(-1, 1, J.shape) * rel)).imag).max()
                for _ in range(16)]

    n_c = n_r = 0
    # complex pairs 0.35 +- i delta, and real pairs 0.35 +- delta (delta as a real split)
    for delta in [0.0, 1e-14, 1e-13, 1e-12, 1e-9, 1e-6, 1e-3, 0.1, -1e-3, -0.1]:
        d = 1j * delta if delta >= 0 else -delta
        for _ in range(20):
            J = _structured_jacobian(rng, [-1.7, -0.9, 1.1, 1.6, 0.35 + d, 0.35 - d])
            if J is None:
                continue
            lam = np.linalg.eigvals(J)
            im = np.abs(lam.imag).max()
            r = CH.fast_eigs_J(J, 1.0)
            assert not (r["real"] and r["cplx"])
            if r["cplx"] and all(i <= thr for i in lapack_imag(J, 4 * 2.0 ** -52) + [im]):
                raise AssertionError(("complex claim where LAPACK is real", delta, im))
            if r["real"]:
                ims = lapack_imag(J, 1e-13) + [im]
                if all(i > thr for i in ims):
                    raise AssertionError(("real claim where LAPACK is complex", delta, im))
                if all(i <= thr for i in ims):
                    n_r += 1
                    assert abs(r["lmin"] - lam.real.min()) < 1e-10 and abs(r["lmax"] - lam.real.max()) < 1e-10
            n_c += r["cplx"]
    print("real claims checked %d, complex claims %d" % (n_r, n_c))
    assert n_r >= 20


def test_mixture_planes_never_claim_a_correction_the_qr_does_not_make():
    """On plane Jacobians of mixtures (complex pairs well above and near the test), a
    complex claim of the fast solve implies the full solve's."""
    rng = np.random.default_rng(21)
    n_c = 0
    for _ in range(300):
        M = to35(mixture_raw(rng, spread=2.0, lam_lo=0.005))
        for f, g in PLANES:
            r = CH.plane_compare(M, f, g)
            if r["fast_cplx"]:
                n_c += 1
                assert r["qr_cplx"], r
    print("complex claims:", n_c)


def test_qr_on_planes_with_subnormal_moments():
    """A Maxwellian whose zero moments are replaced by the smallest subnormal: the Householder
    reduction must not underflow (a column of squares below the normal range), so the full
    solve converges to the Maxwellian's plane extremes u +- sqrt(3) sigma."""
    M = _gauss(0.1, (0.0, 0.0, 0.0), np.eye(3))
    rng = np.random.default_rng(7)
    for _ in range(8):
        Mp = M.copy()
        flip = rng.random(35) < 0.5
        Mp[flip] = np.nextafter(Mp[flip], np.inf)
        for f, g in PLANES:
            r = CH.plane_compare(Mp, f, g)
            assert not r["qr_cplx"], r
            assert abs(r["qr"][0] + np.sqrt(3)) < 1e-12 and abs(r["qr"][1] - np.sqrt(3)) < 1e-12, r
