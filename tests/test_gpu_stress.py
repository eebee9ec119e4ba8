"""The closure + wave-speed stress batch (2^22 Gaussian-mixture cells, every tenth within
1e-8 of the realizability boundary) through hyq_wavespeed_batch.

A subset is compared with the oracle cell by cell at the FP64 bar of parity_util (1e-12
field-scaled, or 32 x the implementations' own spread under one-ulp input changes), with the cells
where that spread exceeds 1e-8 counted and bounded -- the near-boundary cells where the
reference's own decisions (signs of vanishing minors, the 1000 eps complex test) hinge on
rounding.  The full batch is checked through
properties that do not depend on its size: run-to-run bitwise determinism, finite wave
speeds on all but the few cells where the reference's NaN guard applies, ordered wave
speeds, and density kept by the projection."""
import time

import numpy as np
import pytest

import oracle as O
import parity_util as PU

pytestmark = pytest.mark.gpu


def _run(Md, n, precision=0):
    from gpu_util import empty_dev
    from hyqmom_amd import _lib as L
    Mr = empty_dev((35, n))
    F = empty_dev((105, n))
    vp = empty_dev((6, n))
    fl = empty_dev((n,), "int32")
    p = F.data_ptr()
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L.check(L.lib().hyq_wavespeed_batch(Md.data_ptr(), n, 1.5, 0, 1, precision, Mr.data_ptr(), p,
                                        p + 35 * n * 8, p + 70 * n * 8, vp.data_ptr(), fl.data_ptr()),
            "hyq_wavespeed_batch")
    el = time.perf_counter() - t0
    return Mr, F, vp, fl, el


def test_stress_subset_matches_oracle():
    from gpu_util import to_dev, to_host
    from hyqmom_amd.workloads import stress_batch
    n = 4096
    M = stress_batch(n, workers=1)
    Mr, F, vp, fl, _ = _run(to_dev(np.ascontiguousarray(M.T)), n)
    g = np.hstack([to_host(Mr).T, to_host(F).T, to_host(vp).T])
    gf = to_host(fl).view(np.uint32)

    def ora(X):
        Mr_, Fx_, Fy_, Fz_, vp_, _ = O.pass1_batch(X, 0, 1.5, True)
        return np.hstack([Mr_, Fx_, Fy_, Fz_, vp_])

    o = ora(M)
    ofl = O.pass1_batch(M, 0, 1.5, True)[5]
    branch = (1 << 12) - 1  # branch bits, without the fallback / iteration fields
    near = np.arange(n) % 10 == 0
    match = (gf & branch) == (ofl & branch)
    print("flag match: all %.4f, near-boundary %.4f, interior %.4f"
          % (match.mean(), match[near].mean(), match[~near].mean()))
    def dev(X):
        Mr_, F_, vp_, _, _ = _run(to_dev(np.ascontiguousarray(X.T)), len(X))
        return np.hstack([to_host(Mr_).T, to_host(F_).T, to_host(vp_).T])

    sens = PU.sensitivity2(ora, dev, M, o, g, np.random.default_rng(21), nrep=6)
    # the interior cells include mixtures with covariance eigenvalues down to 1e-4
    PU.check(g[~near], o[~near], sens[~near], max_chaotic=0.03, label="stress interior")
    PU.check(g[near], o[near], sens[near], max_chaotic=0.05, label="stress near-boundary")


@pytest.mark.slow
def test_stress_full_batch_properties():
    import torch
    from gpu_util import to_dev, to_host
    from hyqmom_amd.workloads import stress_batch
    n = 1 << 22
    t0 = time.time()
    M = stress_batch(n, workers=8)
    print("generated %d cells in %.1fs" % (n, time.time() - t0), flush=True)
    Md = to_dev(np.ascontiguousarray(M.T))
    del M
    Mr1, F1, vp1, fl1, el1 = _run(Md, n)
    Mr2, F2, vp2, fl2, el2 = _run(Md, n)
    print("hyq_wavespeed_batch 2^22 cells: %.1f ms, %.1f ms -> %.3e cells/s" % (el1 * 1e3, el2 * 1e3, n / el2))
    for a, b in ((Mr1, Mr2), (F1, F2), (vp1, vp2), (fl1, fl2)):
        # bitwise, NaN payloads included
        iv = torch.int64 if a.dtype == torch.float64 else torch.int32
        assert torch.equal(a.view(iv), b.view(iv))
    fl = to_host(fl1).view(np.uint32)
    nonfin = (fl >> 10) & 1 == 1
    vp = to_host(vp1)
    # NaN speeds without the z flag are the reference's own NaN guard on an x/y Jacobian
    # (compute_jacobian_eigenvalues.jl:25-29); the oracle gives them on a few cells per 10^5
    # of this batch too
    fin = np.isfinite(vp).all(axis=0)
    assert fin.mean() > 0.999
    assert (vp[0::2, fin] <= vp[1::2, fin]).all()
    rho_in = to_host(Md[0])
    rho_out = to_host(Mr1[0])
    assert np.allclose(rho_out[~nonfin], rho_in[~nonfin], rtol=1e-14, atol=0)
    print("non-finite z Jacobian cells: %d, cells with a NaN speed: %d; flag histogram: %s"
          % (nonfin.sum(), (~fin).sum(), [int(((fl >> k) & 1).sum()) for k in range(12)]))
