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
