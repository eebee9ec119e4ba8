"""run_simulation on the device against the reference's own golden runs (test/goldenfiles/
mpi_1rank_{small,medium}.bin as fixtures): Nz = 1 crossing jets, Kn = 1, Ma = 0, CFL = 0.5.
FP64 to rounding (1e-12 of each moment's scale, the same step count and final time); the
mixed modes within the reference's own acceptance for these files (test_mpi.jl:35,
1e-6 absolute)."""
import numpy as np
import pytest

from test_oracle_golden_runs import golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["small", "medium"])
@pytest.mark.parametrize("precision", ["fp64", "mixed_newton", "mixed_learned"])
def test_run_simulation_matches_golden_run(name, precision):
    import hyqmom_amd as H
    Np, tmax, tf, steps, G = golden(name)
    r = H.run_simulation(Np=Np, tmax=tmax, Kn=1.0, Ma=0.0, flag2D=0, CFL=0.5, verbose=False,
                         precision=precision)
    M = r["M"]
    assert M.shape == G.shape and r["Nz"] == 1
    assert r["time_steps"] == steps and abs(r["final_time"] - tf) <= 1e-15
    absmax = np.abs(M - G).max()
    if precision == "fp64":
        scale = np.maximum(np.abs(G).max(axis=(0, 1, 2)), 1e-300)
        err = (np.abs(M - G).max(axis=(0, 1, 2)) / scale).max()
        assert absmax < 1e-13 and err < 1e-12, (absmax, err)
    else:
        assert absmax < 1e-6, absmax


@pytest.mark.parametrize("Ma", [0.5, 1.0])
def test_run_simulation_2d_flag_matches_oracle(Ma):
    """flag2D = 1 (the 2D closure and realizability of the reference's 2D mode) at Nz = 1:
    run_simulation on the device against the oracle's step loop, 7 steps to t = 0.05."""
    import math

    import oracle as O
    import hyqmom_amd as H
    from hyqmom_amd.runner import hardcoded_jets
    Np, tmax = 20, 0.05
    r = H.run_simulation(Np=Np, tmax=tmax, Kn=1.0, Ma=Ma, flag2D=1, CFL=0.5, verbose=False)
    M0 = hardcoded_jets(Np, Np, 1, 1.0, 0.01, 1.0, Ma)
    dx = 1.0 / Np
    Mo, to, nno, _, rc = O.run(M0, Np, Np, 1, 1.0, Ma, 1, 0.5, 0.5 * dx, dx, dx, 1.0, True, tmax,
                               math.ceil(tmax / (0.5 * dx)) + 100000)
    assert rc == 0 and r["time_steps"] == nno and abs(r["final_time"] - to) <= 1e-15
    scale = np.maximum(np.abs(Mo).max(axis=(0, 1, 2)), 1e-300)
    err = (np.abs(r["M"] - Mo).max(axis=(0, 1, 2)) / scale).max()
    assert err < 1e-11, err
