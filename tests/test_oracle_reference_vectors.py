"""The oracle against the reference's own per-cell functions.

tests/golden/reference_vectors.npz holds input/output vectors of the reference's Julia
functions (Flux_closure35_and_realizable_3D with flag2D 0 and 1, eigenvalues6_hyperbolic_3D
for x and y, eigenvalues6z_hyperbolic_3D with positivity on and off, realizable_3D with
faces inputs, check2D, collision35 with positivity on and off, closure_and_eigenvalues,
Moments5_3D), evaluated through a statement-by-statement translation of the reference's
sources (tests/golden/make_reference_vectors.py).  The cells were picked so that every
branch the oracle flags is covered: univariate clamp, edge/corner, faces, flag220, 2D
roots, 4x4-minor roots, x/y/z eigenvalue corrections, the z negative-variance
re-projection and the density clamp.

Bar, per vector (errors relative to the largest output magnitude of the vector).  The
sensitivity of a vector is the larger of the reference's and the oracle's: how far their
outputs move when the input is perturbed by 1 to 4 ulps.
- well-conditioned (sensitivity <= 1e-9): |oracle - reference| <= max(1e-12, 4 x sensitivity);
- ill-conditioned (sensitivity <= 1e-2): the difference is within 4 x the sensitivity (the
  reference's own spread at rounding level), or, for a branch decided on a difference at
  rounding level, the two agree to 1e-6 once the input moves by at most 4 ulps (the oracle's
  output is one the reference gives near x, or the reference's one the oracle gives);
- chaotic (larger, or non-finite nearby: degenerate states such as a negative variance or
  density, where both move by O(1) at rounding level): counted, at most 30% of a group.
A vector matched to max(1e-12, 4 x min(sensitivity, 1e-9)) pins its branches whatever its
class; every flagged branch must have at least four such vectors.
"""
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
V = np.load(os.path.join(HERE, "golden", "reference_vectors.npz"))
GROUPS = sorted(set(k.rsplit("_", 1)[0] for k in V.files if k.endswith("_x")))
F_BITS = dict(UNIV=1 << 0, EDGE=1 << 1, FACES=1 << 2, F220=1 << 3, R2D=1 << 4, R4=1 << 5,
              ECX=1 << 6, ECY=1 << 7, ECZ=1 << 8, ZNEG=1 << 9, RHO=1 << 11)


def oracle_eval(group, x):
    if group == "closure":
        return np.concatenate(O.flux_closure35(x, 0, 1.5)[:4])
    if group == "closure2d":
        return np.concatenate(O.flux_closure35(x, 1, 1.0)[:4])
    if group in ("eig1", "eig2"):
        r = O.eigenvalues6(x, int(group[3]), 0, 1.5)
        return np.concatenate([[r[0], r[1]], r[2]])
    if group.startswith("eigz"):
        r = O.eigenvalues6z(x, 0, 1.5, group.endswith("1"))
        return np.concatenate([[r[0], r[1]], r[2]])
    if group == "real3d":
        o, f = O.realizable_3D(x)
        return np.concatenate([o, [f]])
    if group == "check2d":
        return O.check2D(x)
    if group.startswith("bgk"):
        return O.collision35(x, 3.1e-3, 0.01, group.endswith("1"))
    if group == "closure1d":
        return np.array(O.closure_and_eigenvalues(x))
    if group == "moments5":
        return np.concatenate(O.moments5_3d(x))
    raise KeyError(group)


def oracle_sensitivity(group, x, y, rng, k=24):
    sc = max(np.abs(y).max(), 1e-300)
    w = 0.0
    for t in range(k):
        ulps = 1.0 if t < k // 2 else 4.0
        yp = oracle_eval(group, x * (1 + rng.uniform(-1, 1, x.shape) * ulps * 2.0 ** -52))
        if not np.all(np.isfinite(yp)):
            return np.inf
        w = max(w, np.abs(yp - y).max() / sc)
    return w


def check_group(group):
    """Classify and check every vector of a group; returns (counts, per-flag-bit number of
    vectors matched to max(1e-12, 4 x sensitivity))."""
    X, Y, S, NR, FL = (V[group + s] for s in ("_x", "_y", "_sens", "_near", "_flags"))
    rng = np.random.default_rng(7)
    n = {"well": 0, "ill": 0, "chaotic": 0, "matched": 0}
    per_bit = {k: 0 for k in F_BITS}
    worst = 0.0
    for i in range(len(X)):
        o = oracle_eval(group, X[i])
        sc = max(np.abs(Y[i]).max(), 1e-300)
        err = np.abs(o - Y[i]).max() / sc
        sens = max(S[i], oracle_sensitivity(group, X[i], o, rng))
        if err <= max(1e-12, 4 * min(sens, 1e-9)):
            n["matched"] += 1
            for k, b in F_BITS.items():
                per_bit[k] += bool(int(FL[i]) & b)
        if sens <= 1e-9:
            n["well"] += 1
            worst = max(worst, err)
            assert err <= max(1e-12, 4 * sens), (group, i, err, sens)
        elif sens <= 1e-2:
            n["ill"] += 1
            if err <= 4 * sens:
                continue
            best = min([np.abs(o - r).max() / sc for r in NR[i] if np.all(np.isfinite(r))] + [np.inf])
            for t in range(64):
                xp = X[i] * (1 + rng.uniform(-1, 1, X[i].shape) * (1 + 3 * (t % 2)) * 2.0 ** -52)
                best = min(best, np.abs(oracle_eval(group, xp) - Y[i]).max() / sc)
            assert best <= 1e-6, (group, i, err, sens, best)
        else:
            n["chaotic"] += 1
    print("%s: %s, well-conditioned max err %.2e, matched per flag %s"
          % (group, n, worst, {k: v for k, v in per_bit.items() if v}))
    return n, per_bit, len(X)


@pytest.mark.parametrize("group", GROUPS)
def test_oracle_matches_reference_functions(group):
    n, _, size = check_group(group)
    assert n["well"] >= size // 2
    assert n["chaotic"] <= 0.3 * size


def test_every_branch_is_pinned():
    """Every branch the oracle flags has at least four reference vectors that the oracle
    matches to max(1e-12, 4 x sensitivity) (sensitivity <= 1e-9 there)."""
    tot = {k: 0 for k in F_BITS}
    for g in GROUPS:
        _, per_bit, _ = check_group(g)
        for k, v in per_bit.items():
            tot[k] += v
    print(tot)
    for k, v in tot.items():
        assert v >= 4, (k, v)
