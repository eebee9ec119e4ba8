"""Synthetic workloads of the benchmark configurations.

stress_batch(n, seed): the closure + wave-speed stress batch -- n independent cells, each
the exact 35 raw moments (order <= 4) of a mixture of K in {2,3,4} 3D Gaussians with
Dirichlet(1) weights, means uniform in [-3,3]^3 and covariances Q diag(lam) Q^T (Q Haar,
lam log-uniform in [1e-4, 1]).  Every tenth cell (id % 10 == 0) is pushed to the
realizability boundary: all components share one plane (means projected onto it, one
covariance eigenvector along its normal) with a normal variance chosen so that the
mixture's standardized second-order determinant S2 = 1 + 2 r12 r13 r23 - r12^2 - r13^2 -
r23^2 is about 1e-8.
"""
import itertools
from math import comb

import numpy as np

from .moments import E35


def _pairings(idx):
    if not idx:
        yield []
        return
    a = idx[0]
    for t in range(1, len(idx)):
        rest = idx[1:t] + idx[t + 1:]
        for p in _pairings(rest):
            yield [(a, idx[t])] + p


def gaussian_central(S):
    """Central moments (order <= 4) of N(0, S) for a batch of covariances S [..., 3, 3]:
    dict (i,j,k) -> array (Isserlis)."""
    out = {}
    for (i, j, k) in E35:
        idx = [0] * i + [1] * j + [2] * k
        if len(idx) % 2:
            out[(i, j, k)] = np.zeros(S.shape[:-2])
            continue
        acc = np.zeros(S.shape[:-2]) if idx else np.ones(S.shape[:-2])
        if idx:
            for p in _pairings(idx):
                term = np.ones(S.shape[:-2])
                for a, b in p:
                    term = term * S[..., a, b]
                acc = acc + term
        out[(i, j, k)] = acc
    return out


def raw_from_central(cen, mu):
    """Raw moments from central moments and means [..., 3] (three 1D Taylor shifts)."""
    a = dict(cen)
    for ax in range(3):
        sh = mu[..., ax]
        pw = [np.ones_like(sh), sh, sh * sh, sh * sh * sh, sh * sh * sh * sh]
        for n in range(4, 0, -1):
            for e in E35:
                if e[ax] != n:
                    continue
                acc = 0.0
                for p in range(n + 1):
                    f = list(e)
                    f[ax] = p
                    acc = acc + comb(n, p) * pw[n - p] * a[tuple(f)]
                a[e] = acc
    out = np.empty(mu.shape[:-1] + (35,))
    for n, e in enumerate(E35):
        out[..., n] = a[e]
    return out


def _haar(rng, n):
    A = rng.standard_normal((n, 3, 3))
    Q, R = np.linalg.qr(A)
    Q = Q * np.sign(np.diagonal(R, axis1=1, axis2=2))[:, None, :]
    return Q


def _chunk_job(args):
    seed, lo, hi, every, s2 = args
    rng = np.random.default_rng([seed, lo])
    return lo, _stress_chunk(rng, np.arange(lo, hi), every, s2)


def stress_batch(n, seed=0x5EED0003, boundary_every=10, s2_target=1e-8, chunk=1 << 16,
                 workers=1):
    """Returns M [n, 35] (raw moments; density = sum of weights = 1).  Chunks are seeded by
    (seed, first cell id), so the result does not depend on `workers`."""
    out = np.empty((n, 35))
    jobs = [(seed, lo, min(n, lo + chunk), boundary_every, s2_target) for lo in range(0, n, chunk)]
    if workers > 1:
        from concurrent.futures import ProcessPoolExecutor
        with ProcessPoolExecutor(workers) as ex:
            for lo, r in ex.map(_chunk_job, jobs):
                out[lo:lo + len(r)] = r
    else:
        for j in jobs:
            lo, r = _chunk_job(j)
            out[lo:lo + len(r)] = r
    return out


def _stress_chunk(rng, ids, every, s2_target):
    m = len(ids)
    K = rng.integers(2, 5, m)
    w = rng.dirichlet(np.ones(4), m)
    w = np.where(np.arange(4)[None, :] < K[:, None], w, 0.0)
    w /= w.sum(axis=1, keepdims=True)
    mu = rng.uniform(-3, 3, (m, 4, 3))
    lam = np.exp(rng.uniform(np.log(1e-4), 0.0, (m, 4, 3)))
    Q = _haar(rng, m * 4).reshape(m, 4, 3, 3)
    bnd = (ids % every) == 0
    if bnd.any():
        nrm = rng.standard_normal((m, 3))
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        # project the means onto the plane, put one eigenvector along the normal
        mu_b = mu - np.einsum("mcd,md->mc", mu, nrm)[..., None] * nrm[:, None, :]
        # orthonormal basis (e1, e2, nrm)
        t = rng.standard_normal((m, 3))
        e1 = t - (t * nrm).sum(1, keepdims=True) * nrm
        e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
        e2 = np.cross(nrm, e1)
        ang = rng.uniform(0, 2 * np.pi, (m, 4))
        c, s = np.cos(ang)[..., None], np.sin(ang)[..., None]
        a1 = c * e1[:, None, :] + s * e2[:, None, :]
        a2 = -s * e1[:, None, :] + c * e2[:, None, :]
        Qb = np.stack([a1, a2, np.broadcast_to(nrm[:, None, :], a1.shape)], axis=-1)
        lam_b = lam.copy()
        lam_b[..., 2] = 1e-6
        mu = np.where(bnd[:, None, None], mu_b, mu)
        Q = np.where(bnd[:, None, None, None], Qb, Q)
        lam = np.where(bnd[:, None, None], lam_b, lam)
    Sig = np.einsum("mcij,mcj,mckj->mcik", Q, lam, Q)
    raw = (w[..., None] * raw_from_central(gaussian_central(Sig), mu)).sum(axis=1)
    if bnd.any():
        # rescale the normal variance so that S2 of the mixture is about s2_target
        cov = _cov(raw)
        d = np.sqrt(np.diagonal(cov, axis1=1, axis2=2))
        R = cov / (d[:, :, None] * d[:, None, :])
        S2 = np.linalg.det(R)
        fac = np.clip(s2_target / np.maximum(S2, 1e-300), 1e-8, 1e8)
        b = np.nonzero(bnd)[0]
        lam2 = lam[b].copy()
        lam2[..., 2] = lam2[..., 2] * fac[b, None]
        Sig = np.einsum("mcij,mcj,mckj->mcik", Q[b], lam2, Q[b])
        raw[b] = (w[b, :, None] * raw_from_central(gaussian_central(Sig), mu[b])).sum(axis=1)
    return raw


def _cov(raw):
    m0 = raw[:, 0]
    mean = np.stack([raw[:, 1], raw[:, 5], raw[:, 15]], 1) / m0[:, None]
    second = np.empty((len(raw), 3, 3))
    idx = {(0, 0): 2, (1, 1): 9, (2, 2): 19, (0, 1): 6, (0, 2): 16, (1, 2): 25}
    for (a, b), n in idx.items():
        second[:, a, b] = second[:, b, a] = raw[:, n] / m0
    return second - mean[:, :, None] * mean[:, None, :]


def s2_of(raw):
    cov = _cov(raw)
    d = np.sqrt(np.diagonal(cov, axis1=1, axis2=2))
    return np.linalg.det(cov / (d[:, :, None] * d[:, None, :]))
