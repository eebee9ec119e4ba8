"""bench.py's host-side helpers: the seeded one-ulp perturbations of the initial state that
replace a trajectory that breaks down, and the algorithmic bytes of the update."""
import numpy as np

import bench


def test_perturbation_is_seeded_and_one_ulp():
    M0 = np.asfortranarray(np.random.default_rng(0).random((12, 10, 6, 35)) + 0.5)
    assert bench.perturbed_state(M0, 0) is M0
    a, b = bench.perturbed_state(M0, 3), bench.perturbed_state(M0, 3)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, bench.perturbed_state(M0, 4))
    d = a != M0
    assert not d[..., 1:].any()  # only densities move
    n = int(d[..., 0].sum())
    assert 0 < n <= 100
    assert np.array_equal(a[d], np.nextafter(M0[d], np.inf))


def test_update_bytes():
    # read Mr (35 FP64), three flux blocks (3 x 35) and 6 speeds, write the state (35 FP64)
    assert bench.update_bytes("fp64") == 280 + 840 + 48 + 280
    assert bench.update_bytes("mixed_learned") == 280 + 420 + 24 + 280
