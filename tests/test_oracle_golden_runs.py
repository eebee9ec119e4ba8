"""The oracle's whole step loop against the reference's own golden runs: run_simulation on
one rank (test/create_golden_files.jl), crossing jets at Nz = 1, Kn = 1, Ma = 0, CFL = 0.5,
20x20 to t = 0.1 and 40x40 to t = 0.05 (fixtures under tests/golden/).  The reference's MPI
test accepts 1e-6 absolute; the oracle matches to rounding."""
import math

import numpy as np
import pytest

import oracle as O
from hyqmom_amd.runner import hardcoded_jets


def golden(name):
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "mpi_1rank_%s.npz" % name))
    return int(z["Np"]), float(z["tmax"]), float(z["final_time"]), int(z["steps"]), z["M"]


@pytest.mark.parametrize("name", ["small", "medium"])
def test_oracle_matches_golden_run(name):
    Np, tmax, tf, steps, G = golden(name)
    M0 = hardcoded_jets(Np, Np, 1, 1.0, 0.01, 1.0, 0.0)
    dx = 1.0 / Np
    dtmax = 0.5 * min(dx, dx, 1.0)
    M, t, nn, _, rc = O.run(M0, Np, Np, 1, 1.0, 0.0, 0, 0.5, dtmax, dx, dx, 1.0, True, tmax,
                            math.ceil(tmax / dtmax) + 100000)
    assert rc == 0 and nn == steps and abs(t - tf) <= 1e-15
    scale = np.maximum(np.abs(G).max(axis=(0, 1, 2)), 1e-300)
    err = (np.abs(M - G).max(axis=(0, 1, 2)) / scale).max()
    assert np.abs(M - G).max() < 1e-13 and err < 1e-12, err
