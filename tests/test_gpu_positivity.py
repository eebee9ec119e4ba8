"""POSITIVITY_ENABLED = false (src/HyQMOM.jl:20) on the device against the oracle: the BGK
collision without the temperature floor (collision35.jl:31-33), pass 1 without the z path's
density clamp and re-projection (eigenvalues6z_hyperbolic_3D.jl:27-32,63), and whole steps,
at the FP64 bar of parity_util."""
import numpy as np
import pytest

import oracle as O
import parity_util as PU
from test_gpu_parity import _params, cells, device_steps, pass1_dev, pass1_ora, soa_dev

pytestmark = pytest.mark.gpu


def _lib():
    from hyqmom_amd import _lib
    return _lib


def test_collision_batch_without_positivity():
    """Mixture cells, plus cells whose trace temperature is slightly negative: with the floor
    off the relaxation time is not finite there, on both sides."""
    from gpu_util import empty_dev, to_host
    M = cells(128, 31)
    neg = M[:8].copy()
    for m in neg:  # E[u^2] below E[u]^2 by 1e-9: Theta < 0
        m[2] = m[1] * m[1] / m[0] * (1 - 1e-9)
        m[9] = m[5] * m[5] / m[0] * (1 - 1e-9)
        m[19] = m[15] * m[15] / m[0] * (1 - 1e-9)
    M = np.vstack([M, neg])
    L = _lib()

    def dev(X):
        out = empty_dev((35, len(X)))
        L.check(L.lib().hyq_collision_batch(soa_dev(X).data_ptr(), len(X), 0.002, 0.01, 0, out.data_ptr()),
                "collision")
        return to_host(out).T

    def ora(X):
        return np.array([O.collision35(m, 0.002, 0.01, False) for m in X])

    g, o = dev(M), ora(M)
    assert not np.isfinite(o[-8:]).all()  # the unfloored branch is exercised
    PU.check(g, o, PU.sensitivity2(ora, dev, M, o, g, np.random.default_rng(32)),
             max_chaotic=0.07, label="collision, positivity off")


def test_pass1_batch_without_positivity():
    from hyqmom_amd.workloads import stress_batch
    M = np.vstack([cells(200, 33), stress_batch(1024, workers=1)])

    def dev(X):
        return pass1_dev(X, positivity=0)

    def ora(X):
        return pass1_ora(X, positivity=False)

    g, o = dev(M), ora(M)
    PU.check(g, o, PU.sensitivity2(ora, dev, M, o, g, np.random.default_rng(34)), max_chaotic=0.02,
             label="pass 1, positivity off")


def test_config1_without_positivity():
    import hyqmom_amd as H
    p = dict(_params((20, 20, 10), 1.0, 1.0, 100000, 0.01), positivity=False)
    M0 = np.asfortranarray(H.initial_state(p))
    Mg, t, nn, _ = H.simulation_runner(p)

    def ora(X):
        Y, _, _, _, rc = O.run(X, 20, 20, 10, 1.0, 1.0, 0, 0.7, 1e-2, 0.05, 0.05, 0.1, False, 0.01, 100000)
        assert rc == 0
        return Y.reshape(-1, 35, order="F")

    Mo, to, nno, _, rc = O.run(M0, 20, 20, 10, 1.0, 1.0, 0, 0.7, 1e-2, 0.05, 0.05, 0.1, False, 0.01, 100000)
    assert rc == 0 and nn == nno and t == to
    o, g = Mo.reshape(-1, 35, order="F"), Mg.reshape(-1, 35, order="F")
    sens = PU.sensitivity2(ora, lambda X: device_steps(p, X, nn), M0, o, g, np.random.default_rng(35))
    PU.check(g, o, sens, max_chaotic=0.0, label="config 1, positivity off, %d steps" % nn)


@pytest.mark.parametrize("precision", ["fp64", "mixed_newton"])
def test_small_grid_steps_without_positivity(precision):
    """24x24x12 crossing jets, Ma=1.5, Kn=0.01: three steps, each re-run on the device from
    the oracle's state; FP64 at the bar, the fast mode at 1e-6 of each field."""
    import hyqmom_amd as H
    N = (24, 24, 12)
    p = dict(_params(N, 1.5, 0.01, 3, 10.0), positivity=False, precision=precision)
    A = np.asfortranarray(H.initial_state(p))
    s = H.Solver(p)
    t = 0.0
    for it in range(3):
        s.set_state(A)
        s.t, s.nn = t, it
        s.step()
        g = s.get_state().reshape(-1, 35, order="F")

        def ora(X, t=t):
            Y, _, _, _, rc = O.run(X, *N, 0.01, 1.5, 0, 0.7, 1e-2, 1 / N[0], 1 / N[1], 1 / N[2], False, 10.0, 1, t0=t)
            assert rc == 0
            return Y.reshape(-1, 35, order="F")

        A0 = A
        A, t2, _, _, rc = O.run(A, *N, 0.01, 1.5, 0, 0.7, 1e-2, 1 / N[0], 1 / N[1], 1 / N[2], False, 10.0, 1, t0=t)
        assert rc == 0
        o = A.reshape(-1, 35, order="F")
        pf = dict(p, precision="fp64")
        gf = g if precision == "fp64" else device_steps(pf, A0, 1, t, it)
        # the chaotic cells: those of the FP64 step, oracle's or device's
        sens = PU.sensitivity2(ora, lambda X, t=t, it=it: device_steps(pf, X, 1, t, it), np.asfortranarray(A0),
                               o, gf, np.random.default_rng(40 + it))
        t = t2
        if precision == "fp64":
            PU.check(g, o, sens, label="positivity off, step %d" % it)
        else:
            PU.check(g, o, sens, rtol=1e-6, k=1e6, label="positivity off, %s, step %d" % (precision, it))
    s.close()
