"""The fast (mixed_newton) wave-speed batch against the FP64 batch, through
hyq_wavespeed_batch.

The fast batch runs the fast step's pass-1 kernels: projection, the fast plane solves
(Sturm test, Laguerre extremes), full plane solves (QR) for the planes the fast solve cannot
decide, a fix-up, and the full path for cells with a complex pair or a z-path clamp.  The
projected state and the fluxes come from the same FP64 closure code on the same inputs as
in the FP64 batch, so they agree bit for bit; the wave speeds come from a different root
solver (Laguerre against QR) and agree to 1e-12 (scaled by the largest speed)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HYQ_PREC_FP64, HYQ_PREC_MIXED_NEWTON = 0, 3


def batch(fn, M, precision):
    """fn: the batch entry point; M: [n][35] host cells.  Returns host (Mr, F, vp, flags)."""
    from gpu_util import empty_dev, to_dev, to_host
    from hyqmom_amd import _lib as L
    n = len(M)
    Md = to_dev(np.ascontiguousarray(M.T))
    Mr, F, vp = empty_dev((35, n)), empty_dev((105, n)), empty_dev((6, n))
    fl = empty_dev((n,), "int32")
    p = F.data_ptr()
    L.check(fn(Md.data_ptr(), n, 1.5, 0, 1, precision, Mr.data_ptr(), p, p + 35 * n * 8,
               p + 70 * n * 8, vp.data_ptr(), fl.data_ptr()), "wave-speed batch")
    return to_host(Mr), to_host(F), to_host(vp), to_host(fl).view(np.uint32)


def jets_cells(steps=4, N=(24, 24, 12)):
    """Cells of a developing crossing-jets state (Ma = 1.5, Kn = 0.01): mostly near-
    Maxwellian cells, whose planes have nearly double eigenvalues."""
    import hyqmom_amd as H
    from hyqmom_amd.ensemble import jets_params
    p = jets_params(N, 1.5, 0.01, "fp64")
    s = H.Solver(p)
    s.set_state(H.initial_state(p))
    s.run(10.0, steps)
    M = s.get_state().reshape(-1, 35, order="F")
    s.close()
    return M


def mixed_cells():
    from hyqmom_amd.workloads import stress_batch
    return np.concatenate([jets_cells(), stress_batch(3000, workers=1)])


def test_fast_batch_matches_fp64_batch():
    from hyqmom_amd import _lib as L
    M = mixed_cells()
    fn = L.lib().hyq_wavespeed_batch
    a = batch(fn, M, HYQ_PREC_FP64)
    b = batch(fn, M, HYQ_PREC_MIXED_NEWTON)
    branch = (1 << 12) - 1
    match = (a[3] & branch) == (b[3] & branch)
    fallback = ((b[3] >> 12) & 1).astype(bool)
    print("cells %d, flags matched %d, with full plane solves %d" % (len(M), match.sum(), fallback.sum()))
    assert match.mean() > 0.99
    assert fallback.sum() > 0  # the plane-solve path ran
    assert np.array_equal(a[0][:, match], b[0][:, match], equal_nan=True)
    assert np.array_equal(a[1][:, match], b[1][:, match], equal_nan=True)
    va, vb = a[2][:, match], b[2][:, match]
    fin = np.isfinite(va).all(axis=0)
    assert np.array_equal(np.isfinite(va), np.isfinite(vb))
    scale = np.abs(va[:, fin]).max()
    err = np.abs(va[:, fin] - vb[:, fin]).max() / scale
    print("wave speeds: max scaled difference %.2e" % err)
    assert err <= 1e-12
