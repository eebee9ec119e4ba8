"""Crossing jets at production size: 120x120x60, Ma=1.5, Kn=0.01, CFL=0.7.

One FP64 device step against one oracle step (all host threads), from the initial state and
from the state after 11 FP64 device steps (the jets have crossed), at the FP64 bar of
parity_util.  The mixed modes' step from the same states: per-field (density, velocity,
energy) error against the oracle at most 1e-6 of the field's magnitude, on every cell that
is not chaotic (its FP64 step, oracle's or device's, moves by more than 1e-8 under one-ulp
input changes)."""
import numpy as np
import pytest

import oracle as O
import parity_util as PU

pytestmark = pytest.mark.gpu
N3 = (120, 120, 60)


def _params(precision):
    import hyqmom_amd as H
    bg, jets = H.crossing_jets(1.5)
    return dict(Nx=N3[0], Ny=N3[1], Nz=N3[2], tmax=1e9, Kn=0.01, Ma=1.5, flag2D=0, CFL=0.7,
                Nmom=35, nnmax=10 ** 9, dtmax=1e-2, rhol=1.0, rhor=0.1, T=1.0, r110=0.0,
                r101=0.0, r011=0.0, symmetry_check_interval=0, homogeneous_z=False,
                debug_output=False, use_custom_ic=True, ic_background=bg, ic_jets=jets,
                precision=precision)


def _device_step(M, precision):
    import hyqmom_amd as H
    s = H.Solver(_params(precision))
    s.set_state(M)
    dt = s.step()
    out = s.get_state()
    s.close()
    return out.reshape(-1, 35, order="F"), dt


def _oracle(X):
    Y, t, nn, dts, rc = O.run(X, *N3, 0.01, 1.5, 0, 0.7, 1e-2, 1 / N3[0], 1 / N3[1], 1 / N3[2],
                              True, 1e9, 1)
    assert rc == 0
    return Y.reshape(-1, 35, order="F"), dts[0]


@pytest.fixture(scope="module")
def cases():
    import hyqmom_amd as H
    M0 = np.asfortranarray(H.initial_state(_params("fp64")))
    s = H.Solver(_params("fp64"))
    s.set_state(M0)
    s.run(1e9, 11)
    M11 = np.asfortranarray(s.get_state())
    s.close()
    out = {}
    for name, M in (("initial", M0), ("step 11", M11)):
        o, dt = _oracle(M)
        g, gdt = _device_step(M, "fp64")
        sens = PU.sensitivity2(lambda X: _oracle(X)[0], lambda X: _device_step(X, "fp64")[0], M, o, g,
                               np.random.default_rng(5))
        print("%s: oracle dt %.17g, chaotic cells %d" % (name, dt, int((~(sens <= PU.CHAOTIC)).sum())),
              flush=True)
        out[name] = (M, o, dt, sens, g, gdt)
    return out


@pytest.mark.parametrize("name", ["initial", "step 11"])
def test_fp64_step_matches_oracle(cases, name):
    M, o, dt, sens, g, gdt = cases[name]
    assert abs(gdt - dt) <= 1e-14 * dt, (gdt, dt)
    PU.check(g, o, sens, label="120x120x60 FP64 step from the %s state" % name)


def _fields(X):
    r = X[:, 0]
    return np.stack([r, X[:, 1] / r, X[:, 5] / r, X[:, 15] / r, 0.5 * (X[:, 2] + X[:, 9] + X[:, 19])], axis=1)


@pytest.mark.parametrize("precision", ["mixed", "mixed_newton", "mixed_learned"])
@pytest.mark.parametrize("name", ["initial", "step 11"])
def test_mixed_step_error(cases, name, precision):
    M, o, dt, sens, _, _ = cases[name]
    g, gdt = _device_step(M, precision)
    assert abs(gdt - dt) <= 1e-6 * dt, (gdt, dt)
    fo, fg = _fields(o), _fields(g)
    keep = sens <= PU.CHAOTIC
    scale = np.abs(fo[keep]).max(axis=0)
    err = (np.abs(fg[keep] - fo[keep]) / scale).max(axis=0)
    print("%s from the %s state: per-field max error (rho, u, v, w, E) %s over %d cells (%d chaotic)"
          % (precision, name, ["%.2e" % e for e in err], keep.sum(), (~keep).sum()))
    assert np.all(err <= 1e-6), err
