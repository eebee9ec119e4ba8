"""Closed forms and invariants that pin the oracle beyond the golden files."""
import numpy as np
import pytest

import oracle as O
from _common import E35, gaussian_raw, mixture_raw, to35


def maxwellian(rho, u, T, C=None):
    C = np.eye(3) * T if C is None else np.asarray(C)
    return gaussian_raw(rho, u, C, order=5)


@pytest.mark.parametrize("seed", range(6))
def test_eigvals_matches_lapack(seed):
    rng = np.random.default_rng(seed)
    for _ in range(20):
        A = rng.normal(size=(6, 6))
        w = O.eigvals(A)
        ref = np.linalg.eigvals(A)
        np.testing.assert_allclose(np.sort_complex(w), np.sort_complex(ref), atol=1e-10)
        assert np.sum(np.abs(w.imag) > 0) == np.sum(np.abs(ref.imag) > 0)


@pytest.mark.parametrize("seed", range(4))
def test_jacobian_eigs_match_lapack(seed):
    rng = np.random.default_rng(100 + seed)
    raw = mixture_raw(rng)
    m15 = [raw[(0,0,0)], raw[(0,1,0)], raw[(0,2,0)], raw[(0,3,0)], raw[(0,4,0)], raw[(1,0,0)],
           raw[(1,1,0)], raw[(1,2,0)], raw[(1,3,0)], raw[(2,0,0)], raw[(2,1,0)], raw[(2,2,0)],
           raw[(3,0,0)], raw[(3,1,0)], raw[(4,0,0)]]
    J = O.jacobian6(m15)
    np.testing.assert_allclose(np.sort_complex(O.eigvals(J)), np.sort_complex(np.linalg.eigvals(J)),
                               atol=1e-8)


@pytest.mark.parametrize("u,T", [((0.1, 0.2, 0.3), 1.0), ((-0.7, 0.4, 1.1), 2.5), ((1.2, -0.3, 0.0), 0.3)])
def test_maxwellian_closure_is_gaussian(u, T):
    rho = 0.8
    d = maxwellian(rho, u, T)
    M = to35(d)
    Fx, Fy, Fz, Mr, _ = O.flux_closure35(M, 0, 1.0)
    np.testing.assert_allclose(Mr, M, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(Fx, [d[(e[0] + 1, e[1], e[2])] for e in E35], rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(Fy, [d[(e[0], e[1] + 1, e[2])] for e in E35], rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(Fz, [d[(e[0], e[1], e[2] + 1)] for e in E35], rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("u,T", [((0.1, 0.2, 0.3), 1.0), ((-0.7, 0.4, 1.1), 2.5)])
def test_maxwellian_wave_speeds(u, T):
    M = to35(maxwellian(1.3, u, T))
    s = np.sqrt(T)
    for axis, ui in [(1, u[0]), (2, u[1])]:
        vmin, vmax, _, _ = O.eigenvalues6(M, axis)
        assert abs(vmin - (ui - np.sqrt(3) * s)) < 1e-6
        assert abs(vmax - (ui + np.sqrt(3) * s)) < 1e-6
    vmin, vmax, _, _ = O.eigenvalues6(M, 3)
    # z: WU Jacobian (w speeds) and the VW Jacobian of the reference (v speeds)
    assert abs(vmin - min(u[2], u[1]) + np.sqrt(3) * s) < 1e-6
    assert abs(vmax - max(u[2], u[1]) - np.sqrt(3) * s) < 1e-6


def test_mirror_symmetry():
    rng = np.random.default_rng(7)
    raw = mixture_raw(rng)
    M = to35(raw)
    Mm = np.array([M[n] * (-1) ** E35[n][0] for n in range(35)])
    a = O.eigenvalues6(M, 1, 0, 1.0)
    b = O.eigenvalues6(Mm, 1, 0, 1.0)
    assert abs(a[1] + b[0]) < 1e-10 * max(1, abs(a[1]))
    assert abs(a[0] + b[1]) < 1e-10 * max(1, abs(a[0]))


@pytest.mark.parametrize("u,s", [(0.0, 1.0), (0.3, 0.7), (-1.1, 2.0)])
def test_closure_1d_gaussian(u, s):
    rho = 1.7
    mom = [rho, rho * u, rho * (u**2 + s**2), rho * (u**3 + 3 * u * s**2),
           rho * (u**4 + 6 * u**2 * s**2 + 3 * s**4)]
    Mp, vmin, vmax = O.closure_and_eigenvalues(mom)
    assert abs(vmin - (u - np.sqrt(6) * s)) < 1e-9
    assert abs(vmax - (u + np.sqrt(6) * s)) < 1e-9
    assert abs(Mp - rho * (u**5 + 10 * u**3 * s**2 + 15 * u * s**4)) < 1e-9 * max(1, abs(Mp))


@pytest.mark.parametrize("seed", range(4))
def test_bgk_conserves_and_relaxes(seed):
    rng = np.random.default_rng(seed)
    M = to35(mixture_raw(rng))
    dt, Kn = 0.003, 0.05
    Mo = O.collision35(M, dt, Kn)
    for n in [0, 1, 5, 15]:
        assert abs(Mo[n] - M[n]) < 1e-13 * max(1, abs(M[n]))
    E = lambda X: X[2] + X[9] + X[19]
    assert abs(E(Mo) - E(M)) < 1e-12 * abs(E(M))
    rho = M[0]; um = M[1] / rho; vm = M[5] / rho; wm = M[15] / rho
    Th = ((M[2] / rho - um**2) + (M[9] / rho - vm**2) + (M[19] / rho - wm**2)) / 3
    MG = to35(gaussian_raw(rho, (um, vm, wm), np.eye(3) * Th))
    e = np.exp(-dt / (Kn / (rho * np.sqrt(Th) * 2)))
    np.testing.assert_allclose(Mo, MG - e * (MG - M), rtol=1e-11, atol=1e-12)


def test_hll_constant_and_telescoping():
    rng = np.random.default_rng(3)
    Np = 12
    M = np.tile(rng.uniform(0.5, 1.5, 35), (Np, 1))
    F = np.tile(rng.uniform(-1, 1, 35), (Np, 1))
    out = O.pas_hll(M, F, 0.01, 0.1, -np.ones(Np), np.ones(Np))
    np.testing.assert_allclose(out, M, atol=1e-14)
    # telescoping: interior sum changes only through the boundary fluxes
    M = rng.uniform(0.5, 1.5, (Np, 35)); F = rng.uniform(-1, 1, (Np, 35))
    lo = -rng.uniform(0.5, 2, Np); hi = rng.uniform(0.5, 2, Np)
    out = O.pas_hll(M, F, 0.01, 0.1, lo, hi)
    assert np.all(np.isfinite(out))
    np.testing.assert_allclose(out[0], out[1]); np.testing.assert_allclose(out[-1], out[-2])


def test_realizable3d_gaussian_only_shrinks_correlations():
    rng = np.random.default_rng(11)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    C = Q @ np.diag([0.5, 1.0, 1.5]) @ Q.T
    d = gaussian_raw(1.0, (0, 0, 0), C, 4)
    _, S = O.m2cs4(to35(d))
    s28 = [S[E35.index(tuple(int(c) for c in n[1:]))] for n in
           "S300 S400 S110 S210 S310 S120 S220 S030 S130 S040 S101 S201 S301 S102 S202 S003 S103 S004 S011 S111 S211 S021 S121 S031 S012 S112 S013 S022".split()]
    out, flag = O.realizable_3D(s28)
    assert flag == 0
    s = dict(zip("S300 S400 S110 S210 S310 S120 S220 S030 S130 S040 S101 S201 S301 S102 S202 S003 S103 S004 S011 S111 S211 S021 S121 S031 S012 S112 S013 S022".split(), out))
    s0 = dict(zip("S300 S400 S110 S210 S310 S120 S220 S030 S130 S040 S101 S201 S301 S102 S202 S003 S103 S004 S011 S111 S211 S021 S121 S031 S012 S112 S013 S022".split(), s28))
    for k in ["S110", "S101", "S011"]:
        assert abs(s[k] - 0.9999 * s0[k]) < 1e-15
    for k in ["S400", "S040", "S004", "S220", "S202", "S022", "S310"]:
        assert abs(s[k] - s0[k]) < 1e-12


def test_uniform_maxwellian_dt_rule():
    Nx, Ny, Nz = 4, 3, 2
    u = (0.3, -0.2, 0.1); T = 1.2
    M35 = to35(gaussian_raw(0.5, u, np.eye(3) * T))
    M = np.zeros((Nx, Ny, Nz, 35), order="F")
    M[...] = M35
    dx = 1.0 / Nx; dy = 1.0 / Ny; dz = 1.0 / Nz
    out, t, nn, dts, rc = O.run(M, Nx, Ny, Nz, 1.0, 1.0, 0, 0.7, 1.0, dx, dy, dz, True, 10.0, 1)
    assert rc == 0 and nn == 1
    vmax = max(abs(x) for x in u) + np.sqrt(6 * T)
    assert abs(dts[0] - 0.7 * min(dx, dy, dz) / vmax) < 1e-9
    # a uniform Maxwellian is a steady state of the whole step
    np.testing.assert_allclose(out, M, rtol=1e-10, atol=1e-12)
