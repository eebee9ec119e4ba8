"""The CPU reference of the snapshot file's central and standardized moment fields
(compute_central_field / compute_standardized_field), pinned by Gaussian closed forms."""
import numpy as np

import hyqmom_amd as H
from snapshot_ref import central_standardized


def test_gaussian_cells_standardize_to_isserlis_values():
    C2 = np.array([[1.2, 0.3, 0.2], [0.3, 1.5, 0.4], [0.2, 0.4, 1.8]])
    M = np.empty((2, 1, 1, 35))
    M[0, 0, 0] = H.maxwellian35(0.7, (0.1, -0.2, 0.3), C2)
    M[1, 0, 0] = H.maxwellian35(1.3, (0.0, 0.0, 0.0), np.eye(3))
    C, S = central_standardized(M)
    names = {n: i for i, n in enumerate(H.NAMES35)}
    # central moments of the first cell are the zero-mean Gaussian's (density-normalized)
    c0 = H.maxwellian35(1.0, (0, 0, 0), C2)
    np.testing.assert_allclose(C[0, 0, 0, 2:], c0[2:], rtol=1e-12, atol=1e-12)
    # standardized: unit variances, S400 = 3, S220 = 1 + 2 r^2, S110 = r
    r = 0.3 / np.sqrt(1.2 * 1.5)
    s = S[0, 0, 0]
    assert np.allclose([s[names["M200"]], s[names["M020"]], s[names["M002"]]], 1.0)
    assert np.isclose(s[names["M400"]], 3.0) and np.isclose(s[names["M110"]], r)
    assert np.isclose(s[names["M220"]], 1 + 2 * r * r)
    np.testing.assert_allclose(S[1, 0, 0], H.maxwellian35(1.0, (0, 0, 0), np.eye(3)), atol=1e-14)


def test_writer_streams_each_snapshot(tmp_path):
    """Each snapshot is in the file as soon as it is written (the archive is complete after
    every write), with the reference's key names; close() adds n_snapshots and the grid."""
    from hyqmom_amd.snapshots import SnapshotWriter
    fname = str(tmp_path / "snap.npz")
    w = SnapshotWriter(fname, {"Nx": 2, "Ma": 1.5, "name": "jets"}, 3)
    M = np.arange(2 * 1 * 1 * 35, dtype=float).reshape(2, 1, 1, 35)
    w.write(M, 0.0, 0, M + 1, M + 2)
    with np.load(fname) as f:
        assert set(f.files) == {"meta/snapshot_interval", "meta/params/Nx", "meta/params/Ma",
                                "meta/params/name", "snapshots/000001/M", "snapshots/000001/t",
                                "snapshots/000001/step", "snapshots/000001/S", "snapshots/000001/C"}
        np.testing.assert_array_equal(f["snapshots/000001/C"], M + 1)
    w.write(2 * M, 0.5, 3, M, M)
    w.close({"x": np.arange(3.0), "dx": 0.5})
    with np.load(fname) as f:
        assert int(f["meta/n_snapshots"]) == 2 and int(f["meta/snapshot_interval"]) == 3
        np.testing.assert_array_equal(f["snapshots/000002/M"], 2 * M)
        assert float(f["snapshots/000002/t"]) == 0.5 and int(f["snapshots/000002/step"]) == 3
        np.testing.assert_array_equal(f["grid/x"], np.arange(3.0))
