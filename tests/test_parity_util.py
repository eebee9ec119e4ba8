"""The sensitivity-based parity bar (tests/parity_util.py) on the CPU: the host build of the
device cell code against the oracle meets it, and a mistake planted in the oracle's output
fails it."""
import numpy as np
import pytest

import cellhost as CH
import oracle as O
import parity_util as PU
from _common import mixture_raw, to35


def test_ulp_perturb_moves_nonzero_entries_by_one_ulp():
    rng = np.random.default_rng(0)
    X = rng.standard_normal((50, 35))
    X[:, 3] = 0.0
    Y = PU.ulp_perturb(X, rng)
    assert np.array_equal(Y[:, 3], X[:, 3])
    nz = X != 0
    up, down = np.nextafter(X, np.inf), np.nextafter(X, -np.inf)
    assert np.all((Y[nz] == up[nz]) | (Y[nz] == down[nz]))
    assert (Y[nz] == up[nz]).mean() > 0.3 and (Y[nz] == down[nz]).mean() > 0.3


def _pass1_oracle(M):
    Mr, Fx, Fy, Fz, vp, _ = O.pass1_batch(M, 0, 1.5, True, nthreads=4)
    return np.hstack([Mr, Fx, Fy, Fz, vp])


def _pass1_host(M):
    out = []
    for m in M:
        Mr, Fx, Fy, Fz, vp, _ = CH.pass1(m, 0, 1.5, True)
        out.append(np.concatenate([Mr, Fx, Fy, Fz, vp]))
    return np.array(out)


@pytest.fixture(scope="module")
def batch():
    rng = np.random.default_rng(3)
    M = np.array([to35(mixture_raw(rng, spread=1.5, lam_lo=0.01)) for _ in range(300)])
    o = _pass1_oracle(M)
    sens = PU.sensitivity(_pass1_oracle, M, o, np.random.default_rng(4))
    return M, o, sens


def test_host_build_meets_the_bar(batch):
    M, o, _ = batch
    g = _pass1_host(M)
    PU.check(g, o, PU.sensitivity2(_pass1_oracle, _pass1_host, M, o, g, np.random.default_rng(6)),
             label="host pass 1")


@pytest.mark.parametrize("rel", [1e-10, 1e-6])
def test_planted_mistake_in_the_oracle_fails(batch, rel):
    M, o, sens = batch
    g = _pass1_host(M)
    bad = o.copy()
    # one component of a well-conditioned cell off by rel of its field's scale
    c = int(np.flatnonzero(sens < 1e-13)[5])
    bad[c, 40] += rel * PU.field_scale(o)[40]
    with pytest.raises(AssertionError):
        PU.check(g, bad, sens, label="planted")
