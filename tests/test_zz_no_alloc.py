"""hyq_wavespeed_batch_noalloc (the batch with its scratch in a fixed device array, for
callers that cannot allocate) against hyq_wavespeed_batch: bit for bit, on two calls in a
row, the second with a smaller, different input; and the size limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from test_gpu_fast_batch import HYQ_PREC_FP64, HYQ_PREC_MIXED_NEWTON, batch, mixed_cells


@pytest.mark.parametrize("precision", [HYQ_PREC_MIXED_NEWTON, HYQ_PREC_FP64])
def test_noalloc_matches_allocating_batch(precision):
    from hyqmom_amd import _lib as L
    M = mixed_cells()
    for sub in (M, M[len(M) // 3: len(M) // 3 + 1500]):
        want = batch(L.lib().hyq_wavespeed_batch, sub, precision)
        got = batch(L.lib().hyq_wavespeed_batch_noalloc, sub, precision)
        for w, g in zip(want, got):
            assert np.array_equal(w, g, equal_nan=True)


def test_noalloc_rejects_too_many_cells():
    from gpu_util import empty_dev
    from hyqmom_amd import _lib as L
    n = (1 << 16) + 1
    Md, Mr, F, vp = empty_dev((35, n)), empty_dev((35, n)), empty_dev((105, n)), empty_dev((6, n))
    fl = empty_dev((n,), "int32")
    p = F.data_ptr()
    rc = L.lib().hyq_wavespeed_batch_noalloc(Md.data_ptr(), n, 1.5, 0, 1, HYQ_PREC_MIXED_NEWTON,
                                             Mr.data_ptr(), p, p + 35 * n * 8, p + 70 * n * 8,
                                             vp.data_ptr(), fl.data_ptr())
    assert rc == -1
    assert b"65536" in L.lib().hyq_last_error()
