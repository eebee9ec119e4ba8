"""The reference's API over the device step: simulation_runner's return contract and
run_simulation_with_snapshots' snapshot file (src/simulation_runner.jl:63-759, 994-1011)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _params(tmp_path=None, **kw):
    import hyqmom_amd as H
    bg, jets = H.crossing_jets(1.0)
    # the 20x20x10 crossing jets at Ma = 1, Kn = 1 (stable for these few steps)
    p = dict(Nx=20, Ny=20, Nz=10, tmax=10.0, Kn=1.0, Ma=1.0, flag2D=0, CFL=0.7, Nmom=35, nnmax=5,
             dtmax=1e-2, rhol=1.0, rhor=0.1, T=1.0, r110=0.0, r101=0.0, r011=0.0,
             symmetry_check_interval=0, homogeneous_z=False, debug_output=False,
             use_custom_ic=True, ic_background=bg, ic_jets=jets)
    p.update(kw)
    return p


def test_simulation_runner_contract():
    import hyqmom_amd as H
    M, t, nn, grid = H.simulation_runner(_params())
    assert M.shape == (20, 20, 10, 35) and M.flags["F_CONTIGUOUS"]
    assert nn == 5 and 0 < t < 10.0
    for k in ("x", "y", "z", "dx", "dy", "dz", "xm", "ym", "zm"):
        assert k in grid
    assert len(grid["x"]) == 21 and len(grid["xm"]) == 20 and grid["dz"] == 1.0 / 10
    assert np.isfinite(M).all()
    # the loop condition t < tmax ends the run at tmax exactly (dt = min(..., tmax - t))
    M2, t2, nn2, _ = H.simulation_runner(_params(tmax=0.005, nnmax=100))
    assert t2 == 0.005 and nn2 >= 1


def test_run_simulation_with_snapshots(tmp_path):
    import hyqmom_amd as H
    fname = str(tmp_path / "snap.npz")
    p = _params(snapshot_filename=fname)
    out, grid = H.run_simulation_with_snapshots(p, snapshot_interval=2)
    assert out == fname and os.path.exists(fname)
    z = np.load(fname)
    n = int(z["meta/n_snapshots"])
    steps = [int(z["snapshots/%06d/step" % i]) for i in range(1, n + 1)]
    assert steps == [0, 2, 4, 5]
    M_end, t_end, nn_end, _ = H.simulation_runner(_params())
    assert np.array_equal(z["snapshots/%06d/M" % n], M_end)
    assert float(z["snapshots/%06d/t" % n]) == t_end
    S = z["snapshots/%06d/S" % n]
    assert np.allclose(S[..., 2], 1.0) and np.allclose(S[..., 1], 0.0)
    assert int(z["meta/snapshot_interval"]) == 2 and "grid/xm" in z.files


def test_snapshot_fields_match_reference(tmp_path):
    """hyq_snapshot_fields (device M2CS4_35 per cell) against the CPU reference of
    compute_central_field / compute_standardized_field, on the initial jets and after a few
    steps; the snapshot file holds the device fields of its state."""
    import hyqmom_amd as H
    from snapshot_ref import central_standardized
    p = _params()
    s = H.Solver(p)
    try:
        s.set_state(H.initial_state(p))
        for steps in (0, 3):
            if steps:
                s.run(10.0, steps)
            M = s.get_state()
            C, S = s.snapshot_fields()
            Cr, Sr = central_standardized(M)
            for got, want in ((C, Cr), (S, Sr)):
                # per moment, relative to its largest magnitude or to 1 (moments that vanish
                # exactly in one restatement carry rounding-level values in the other)
                scale = np.maximum(np.abs(want).max(axis=(0, 1, 2)), 1.0)
                err = (np.abs(got - want).max(axis=(0, 1, 2)) / scale).max()
                assert err < 1e-11, (steps, err)
    finally:
        s.close()
    fname = str(tmp_path / "snap.npz")
    H.run_simulation_with_snapshots(_params(snapshot_filename=fname, nnmax=2), snapshot_interval=2)
    z = np.load(fname)
    M = z["snapshots/000002/M"]
    Cr, Sr = central_standardized(M)
    assert np.allclose(z["snapshots/000002/C"], Cr, rtol=1e-10, atol=1e-12)
    assert np.allclose(z["snapshots/000002/S"], Sr, rtol=1e-10, atol=1e-12)
