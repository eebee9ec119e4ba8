"""The CPU reference of the runner's diagnostics, pinned by closed forms: a mirror-symmetric
state has MaxDiff 0, one planted asymmetry of size delta gives delta / (sqrt(2) delta + 1),
and the worst-cell search follows the reference's loop order and skips NaN cells."""
import math

import numpy as np

import diagnostics_ref as D


def _symmetric_state(N=7, Nz=3, seed=0):
    """Moments 1, 3, 5 even and 2, 4 odd under (i, j) -> (N-1-i, N-1-j) on every slice."""
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((N, N, Nz, 35))
    for i in range(N):
        for j in range(N):
            a, b = (i, j), (N - 1 - i, N - 1 - j)
            if a > b:
                continue
            for n in range(5):
                if a == b:
                    if n % 2 == 1:
                        M[i, j, :, n] = 0.0
                    continue
                M[b[0], b[1], :, n] = -M[i, j, :, n] if n % 2 == 1 else M[i, j, :, n]
    return M


def test_symmetric_state_has_zero_maxdiff():
    M = _symmetric_state()
    m, per = D.symmetry_maxdiff(M)
    assert m == 0.0 and np.all(per == 0.0)


def test_planted_asymmetry_closed_form():
    delta = 0.25
    for n in range(5):
        M = _symmetric_state(seed=n + 1)
        M[1, 1, 0, n] += delta  # diagonal cell 1; its mirror is cell 5 (N = 7)
        m, per = D.symmetry_maxdiff(M)
        want = delta / (math.sqrt(2.0) * delta + 1.0)
        assert abs(per[n] - want) < 1e-15, (n, per)
        assert np.all(np.delete(per, n) == 0.0)
        assert abs(m - want) < 1e-15
    # only the k = 1 slice and the first five moments count
    M = _symmetric_state()
    M[1, 1, 1, 0] += 1.0
    M[1, 1, 0, 7] += 1.0
    M[1, 2, 0, 0] += 1.0
    assert D.symmetry_maxdiff(M)[0] == 0.0


def test_vmax_breakdown_order_and_nan():
    Nx, Ny, Nz = 2, 3, 2
    N = Nx * Ny * Nz
    vp = np.zeros((6, N))

    def c(i, j, k):
        return i + Nx * (j + Ny * k)

    vp[0, c(1, 0, 0)] = -2.0   # x min
    vp[3, c(0, 2, 0)] = 3.0    # y max
    vp[5, c(1, 1, 1)] = 3.0    # z max, the same speed at a later k
    vp[4, c(0, 1, 0)] = -1.5
    vdir, worst = D.vmax_breakdown(vp, Nx, Ny, Nz)
    assert vdir == (2.0, 3.0, 3.0)
    assert worst == (0, 2, 0)  # k first, then i, then j
    vp[2, c(0, 0, 0)] = 3.0   # same speed, earlier j at the same k and i
    assert D.vmax_breakdown(vp, Nx, Ny, Nz)[1] == (0, 0, 0)
    vp[1, c(1, 2, 0)] = np.nan
    vdir, worst = D.vmax_breakdown(vp, Nx, Ny, Nz)
    assert math.isnan(vdir[0]) and vdir[1:] == (3.0, 3.0)
    assert worst == (0, 0, 0)
    vp[:, :] = np.nan
    assert D.vmax_breakdown(vp, Nx, Ny, Nz)[1] == (-1, -1, -1)
