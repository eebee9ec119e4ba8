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
