"""The FP64 parity bar, with the knife-edge exemption defined by each implementation's own
conditioning.

A cell's sensitivity to an implementation is how far that implementation's output for it
moves (field-scaled: each output component divided by its largest magnitude over the batch)
when every nonzero input value moves by one ulp, up or down at random (zeros stay zero: they
are exact symmetries of the inputs, and a subnormal there is no rounding-level change),
taking the largest change over a few such perturbations.  The cell's sensitivity is the
larger of the oracle's and the device's: two correct FP64 implementations that round
differently differ by about that much.  Then, per cell:
- sensitivity <= 1e-8: the device agrees with the oracle to max(1e-12, 32 x sensitivity)
  (the spread under rounding, and 1e-12 where that spread is smaller);
- sensitivity > 1e-8 or non-finite (a branch of the reference decided by a difference at
  rounding level: a vanishing minor, a discriminant that is exactly zero at a symmetric
  state, the 1000 eps complex test, a clamp): chaotic, counted, and bounded as a fraction of
  the batch.
No exemption depends on the device/oracle difference itself: a wrong device formula leaves
the device's own sensitivity small and its difference large.
"""
import numpy as np

CHAOTIC = 1e-8
K_SENS = 32.0


def ulp_perturb(X, rng):
    """Every nonzero entry of X moved by one ulp, up or down at random; zeros kept."""
    Y = np.array(X, dtype=np.float64, copy=True)
    nz = Y != 0
    up = rng.random(Y.shape) < 0.5
    Y[nz & up] = np.nextafter(Y[nz & up], np.inf)
    Y[nz & ~up] = np.nextafter(Y[nz & ~up], -np.inf)
    return Y


def field_scale(o):
    """Per output component: the largest finite magnitude over the cells ([n, k] -> [k])."""
    a = np.where(np.isfinite(o), np.abs(o), 0.0)
    s = a.max(axis=0)
    return np.where(s > 0, s, 1e-300)


def cell_err(a, b, scale):
    """Per cell: the largest field-scaled difference ([n, k] -> [n]); equal values, NaN
    against NaN included, differ by 0, a finite value against a non-finite one by inf."""
    with np.errstate(invalid="ignore", over="ignore"):
        d = np.abs(a - b) / scale
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    d = np.where(same, 0.0, d)
    d = np.where(np.isnan(d), np.inf, d)
    return d.max(axis=1)


def sensitivity(f, X, o, rng, nrep=3):
    """Per cell: the largest field-scaled change of f's output ([n, k]) over nrep one-ulp
    perturbations of its input X (any shape that f takes)."""
    scale = field_scale(o)
    s = np.zeros(len(o))
    for _ in range(nrep):
        s = np.maximum(s, cell_err(f(ulp_perturb(X, rng)), o, scale))
    return s


def sensitivity2(ora, dev, X, o, g, rng, nrep=3):
    """Per cell: the larger of the oracle's sensitivity (ora, output o) and the device's (dev,
    output g) on the same perturbations of X."""
    scale = field_scale(o)
    s = np.zeros(len(o))
    for _ in range(nrep):
        Xp = ulp_perturb(X, rng)
        s = np.maximum(s, cell_err(ora(Xp), o, scale))
        s = np.maximum(s, cell_err(dev(Xp), g, scale))
    return s


def check(g, o, sens, rtol=1e-12, k=K_SENS, max_chaotic=0.01, label=""):
    """Asserts the bar on device output g against oracle output o ([n, k] each) with the
    oracle's per-cell sensitivity sens; returns (chaotic cells, largest error on the others,
    the largest error / allowed ratio)."""
    scale = field_scale(o)
    e = cell_err(g, o, scale)
    chaotic = ~(sens <= CHAOTIC)
    lim = np.maximum(rtol, k * sens)
    ok = ~chaotic
    ratio = (e[ok] / lim[ok]).max() if ok.any() else 0.0
    worst = e[ok].max() if ok.any() else 0.0
    print("%s: %d cells, chaotic %d (%.3f%%), others: max error %.2e, max error/allowed %.3f, "
          "cells with sensitivity above 1e-12: %d"
          % (label, len(e), chaotic.sum(), 100.0 * chaotic.mean(), worst, ratio,
             int((ok & (sens > 1e-12)).sum())))
    bad = np.flatnonzero(ok & ~(e <= lim))
    assert len(bad) == 0, (label, bad[:10], e[bad[:10]], sens[bad[:10]])
    assert chaotic.mean() <= max_chaotic, (label, chaotic.sum())
    return int(chaotic.sum()), worst, ratio
