"""The device cell numerics (hyqmom.jl_amd/csrc, compiled for the host by tests/native)
against the oracle, on CPU.  Branch flags must match; on matched cells the values agree
to the FP64 bar."""
import numpy as np
import pytest

import cellhost as C
import oracle as O
from _common import gaussian_raw, mixture_raw, to35


def _err(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12 * max(1.0, np.abs(b).max())))


@pytest.mark.parametrize("seed", range(3))
def test_closure_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    worst, flips = 0.0, 0
    for _ in range(60):
        M = to35(mixture_raw(rng, spread=1.5, lam_lo=0.01))
        o = O.flux_closure35(M, 0, 1.5)
        c = C.flux_closure35(M, 0, 1.5)
        if o[4] != c[4]:
            flips += 1
            continue
        worst = max(worst, *[_err(x, y) for x, y in zip(c[:4], o[:4])])
    assert flips <= 1
    assert worst < 1e-11, worst


@pytest.mark.parametrize("seed", range(3))
def test_pass1_matches_oracle(seed):
    rng = np.random.default_rng(10 + seed)
    worst, flips = 0.0, 0
    for _ in range(40):
        M = to35(mixture_raw(rng, spread=1.5, lam_lo=0.01))
        o = O.pass1_batch(M[None, :], 0, 1.5, True, 1)
        c = C.pass1(M, 0, 1.5)
        if o[5][0] != c[5]:
            flips += 1
            continue
        for k in range(5):
            worst = max(worst, _err(c[k], o[k][0]))
    assert flips <= 1
    assert worst < 1e-10, worst


def test_collision_and_1d():
    rng = np.random.default_rng(5)
    for _ in range(20):
        M = to35(mixture_raw(rng))
        assert _err(C.collision35(M, 0.003, 0.02), O.collision35(M, 0.003, 0.02)) < 1e-12
        mom = M[[0, 1, 2, 3, 4]]
        _, a, b = O.closure_and_eigenvalues(mom)
        a2, b2 = C.closure1d(mom)
        assert abs(a - a2) < 1e-9 * max(1, abs(a)) and abs(b - b2) < 1e-9 * max(1, abs(b))


def test_maxwellian_eig_closed_form():
    d = gaussian_raw(0.7, (0.3, -0.4, 0.2), np.eye(3) * 1.3, 4)
    vmin, vmax, _, _ = C.eigenvalues6(to35(d), 1)
    assert abs(vmin - (0.3 - np.sqrt(3 * 1.3))) < 1e-6
    assert abs(vmax - (0.3 + np.sqrt(3 * 1.3))) < 1e-6


def _field(shape, seed, mixtures=True):
    """A field of Gaussian cells, optionally with a few non-Gaussian mixture cells."""
    rng = np.random.default_rng(seed)
    Nx, Ny, Nz = shape
    M = np.empty(shape + (35,), order="F")
    for i in range(Nx):
        for j in range(Ny):
            for k in range(Nz):
                u = (0.3 * np.sin(i), 0.2 * np.cos(j), 0.1 * k)
                M[i, j, k] = to35(gaussian_raw(0.5 + 0.1 * ((i + j + k) % 3), u, np.eye(3) * (1 + 0.1 * i), 4))
    for _ in range(max(1, Nx * Ny * Nz // 10) if mixtures else 0):
        i, j, k = rng.integers(0, Nx), rng.integers(0, Ny), rng.integers(0, Nz)
        M[i, j, k] = to35(mixture_raw(rng, spread=1.0, lam_lo=0.1))
    return M


@pytest.mark.parametrize("shape", [(6, 5, 4), (7, 3, 1), (2, 5, 3), (1, 4, 3)])
def test_host_step_matches_oracle(shape):
    """The device step's update (per-cell HLL face data, BC copies, Nz == 1 and 1- and
    2-cell lines) and passes, emulated on the host, against the oracle's step.  Gaussian
    cells only: no cell sits on a realizability or eigenvalue knife edge, so every cell
    must agree."""
    M = _field(shape, 3, mixtures=False)
    Nx, Ny, Nz = shape
    got, dt = C.host_step(M, 0.05, 1.0)
    ref, t, nn, dts, rc = O.run(M, Nx, Ny, Nz, 0.05, 1.0, 0, 0.7, 1e-2, 1.0 / Nx, 1.0 / Ny, 1.0 / Nz,
                                True, 1e9, 1)
    assert rc == 0 and abs(dt - t) <= 1e-15 * t
    scale = np.abs(ref).max(axis=(0, 1, 2))
    err = (np.abs(got - ref) / np.maximum(scale, 1e-300)).max()
    assert err < 1e-12, err
