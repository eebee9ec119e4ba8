"""Write input/output vectors of the reference's symbolic-toolbox functions.

The outputs come from the reference formulas themselves (src/autogen/*.jl translated to
oracle/_ref/autogen_ref.py by oracle/ref_autogen.py; that file is never committed).  The
inputs are standardized / raw moments of seeded random 3D Gaussian mixtures (realizable),
plus random non-realizable argument sets for the Cardano root functions.
Writes tests/golden/autogen_vectors.json.
Usage: python oracle/ref_autogen.py && python tests/golden/make_autogen_vectors.py
"""
import itertools
import json
import os
import sys
from math import comb

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle", "_ref"))
import autogen_ref as A  # noqa: E402

E35 = [(0,0,0),(1,0,0),(2,0,0),(3,0,0),(4,0,0),(0,1,0),(1,1,0),(2,1,0),(3,1,0),(0,2,0),
       (1,2,0),(2,2,0),(0,3,0),(1,3,0),(0,4,0),(0,0,1),(1,0,1),(2,0,1),(3,0,1),(0,0,2),
       (1,0,2),(2,0,2),(0,0,3),(1,0,3),(0,0,4),(0,1,1),(1,1,1),(2,1,1),(0,2,1),(1,2,1),
       (0,3,1),(0,1,2),(1,1,2),(0,1,3),(0,2,2)]
N28 = ("S300 S400 S110 S210 S310 S120 S220 S030 S130 S040 S101 S201 S301 S102 S202 S003 "
       "S103 S004 S011 S111 S211 S021 S121 S031 S012 S112 S013 S022").split()


def key(n):
    return tuple(int(c) for c in n[1:])


def gauss_raw(m, C, e):
    """E[x^i y^j z^k] for N(m, C) by summing over perfect matchings (Isserlis)."""
    idx = [0] * e[0] + [1] * e[1] + [2] * e[2]
    n = len(idx)

    def pm(l):
        if not l:
            return 1.0
        a = l[0]
        return sum(C[a][l[t]] * pm(l[1:t] + l[t + 1:]) for t in range(1, len(l)))
    tot = 0.0
    for mask in range(1 << n):
        sel = [idx[t] for t in range(n) if mask >> t & 1]
        if len(sel) % 2:
            continue
        p = 1.0
        for t in range(n):
            if not mask >> t & 1:
                p *= m[idx[t]]
        tot += p * pm(sel)
    return tot


def mixture(rng, order=5):
    K = int(rng.integers(2, 5))
    w = rng.dirichlet(np.ones(K))
    mu = rng.uniform(-1, 1, (K, 3))
    covs = []
    for _ in range(K):
        Q, _r = np.linalg.qr(rng.normal(size=(3, 3)))
        lam = np.exp(rng.uniform(np.log(0.05), 0, 3))
        covs.append(Q @ np.diag(lam) @ Q.T)
    rho = rng.uniform(0.2, 2.0)
    raw = {}
    for i in range(order + 1):
        for j in range(order + 1 - i):
            for k in range(order + 1 - i - j):
                raw[(i, j, k)] = rho * sum(w[c] * gauss_raw(mu[c], covs[c], (i, j, k)) for c in range(K))
    return raw


def central(raw, order):
    m0 = raw[(0, 0, 0)]
    mu = [raw[(1, 0, 0)] / m0, raw[(0, 1, 0)] / m0, raw[(0, 0, 1)] / m0]
    cen = {}
    for (i, j, k) in raw:
        if i + j + k > order:
            continue
        s = 0.0
        for p in range(i + 1):
            for q in range(j + 1):
                for r in range(k + 1):
                    s += (comb(i, p) * comb(j, q) * comb(k, r) * (-mu[0]) ** (i - p) *
                          (-mu[1]) ** (j - q) * (-mu[2]) ** (k - r) * raw[(p, q, r)] / m0)
        cen[(i, j, k)] = s
    return mu, cen


def main():
    rng = np.random.default_rng(0x5EED0A)
    out = {k: [] for k in ["M4toC4_3D", "C4toM4_3D", "C5toM5_3D", "delta2star3D", "lower_bound_S220",
                           "bound_minor1", "delta2starchol_L3", "rootsR", "rootsR_X_Y", "jacobian6"]}
    for case in range(12):
        raw = mixture(rng)
        m0 = raw[(0, 0, 0)]
        mu, cen = central(raw, 5)
        sd = np.sqrt([cen[(2, 0, 0)], cen[(0, 2, 0)], cen[(0, 0, 2)]])
        S = {k: cen[k] / (sd[0] ** k[0] * sd[1] ** k[1] * sd[2] ** k[2]) for k in cen}
        M35 = [raw[e] for e in E35]
        C = np.array(A.M4toC4_3D(*M35), dtype=float)
        out["M4toC4_3D"].append({"in": M35, "out": C.ravel(order="F").tolist()})
        cargs = [cen[k] for k in [(2,0,0),(1,1,0),(1,0,1),(0,2,0),(0,1,1),(0,0,2),(3,0,0),(2,1,0),
                                   (2,0,1),(1,2,0),(1,1,1),(1,0,2),(0,3,0),(0,2,1),(0,1,2),(0,0,3),
                                   (4,0,0),(3,1,0),(3,0,1),(2,2,0),(2,1,1),(2,0,2),(1,3,0),(1,2,1),
                                   (1,1,2),(1,0,3),(0,4,0),(0,3,1),(0,2,2),(0,1,3),(0,0,4)]]
        R4 = np.array(A.C4toM4_3D(m0, *mu, *cargs), dtype=float)
        out["C4toM4_3D"].append({"in": [m0] + mu + cargs, "out": R4.ravel(order="F").tolist()})
        c5 = [cen[k] for k in [(5,0,0),(4,1,0),(3,2,0),(2,3,0),(1,4,0),(4,0,1),(3,0,2),(2,0,3),
                                (1,0,4),(3,1,1),(2,2,1),(1,3,1),(2,1,2),(1,1,3),(1,2,2),(0,5,0),
                                (0,4,1),(0,3,2),(0,2,3),(0,1,4),(0,0,5)]]
        R5 = np.array(A.C5toM5_3D(m0, *mu, *cargs, *c5), dtype=float)
        out["C5toM5_3D"].append({"in": [m0] + mu + cargs + c5, "out": R5.ravel(order="F").tolist()})
        s28 = [S[key(n)] for n in N28]
        E = np.array(A.delta2star3D(*s28), dtype=float)
        out["delta2star3D"].append({"in": s28, "out": E.ravel().tolist()})
        lbin = [S[key(n)] for n in "S011 S012 S021 S101 S102 S110 S111 S120 S201 S210".split()]
        out["lower_bound_S220"].append({"in": s28, "out": np.array(A.lower_bound_S220(*lbin), dtype=float).tolist()})
        bmin = [S[key(n)] for n in "S003 S011 S012 S021 S030 S101 S102 S110 S111 S120 S201 S210 S300".split()]
        out["bound_minor1"].append({"in": s28, "out": np.array(A.bound_minor1(*bmin), dtype=float).ravel().tolist()})
        l3in = [S[(0,0,0)] * 0 + S[(0,3,0)], S[(0,4,0)], S[(1,1,0)], S[(1,2,0)], S[(1,3,0)], S[(2,1,0)],
                S[(2,2,0)], S[(3,0,0)], S[(3,1,0)], S[(4,0,0)]]
        out["delta2starchol_L3"].append({"in": l3in, "out": float(A.delta2starchol_L3(*l3in))})
        for plane in [(0, 1), (0, 2), (2, 1)]:
            def m2(i, j):
                e = [0, 0, 0]; e[plane[0]] += i; e[plane[1]] += j
                return raw[tuple(e)]
            m15 = [m2(0,0), m2(0,1), m2(0,2), m2(0,3), m2(0,4), m2(1,0), m2(1,1), m2(1,2), m2(1,3),
                   m2(2,0), m2(2,1), m2(2,2), m2(3,0), m2(3,1), m2(4,0)]
            J = np.array(A.jacobian6(*m15), dtype=float)
            out["jacobian6"].append({"in": m15, "out": J.ravel().tolist()})
    # Cardano roots on random argument sets: keep cases with a real+complex pair (finite)
    n_xy = n_r = 0
    while n_xy < 12 or n_r < 12:
        e = rng.normal(size=9).tolist()
        R = np.array(A.rootsR_X_Y(*e), dtype=complex)
        if n_xy < 12 and np.all(np.isfinite(R)):
            out["rootsR_X_Y"].append({"in": e, "re": R.real.tolist(), "im": R.imag.tolist()}); n_xy += 1
        D1, H02, H20 = rng.uniform(0.01, 1, 3)
        s03, s12, s13, s21, s30, s31 = rng.normal(size=6)
        s11 = float(np.sqrt(1 - D1) * rng.choice([-1, 1]))
        a10 = [float(D1), float(H02), float(H20), s03, s11, s12, s13, s21, s30, s31]
        R = np.array(A.rootsR(*a10), dtype=complex)
        if n_r < 12 and np.all(np.isfinite(R)):
            out["rootsR"].append({"in": a10, "re": R.real.tolist(), "im": R.imag.tolist()}); n_r += 1
    # three-real-root cases: the reference returns NaN (real sqrt of a negative discriminant)
    nan_cases = []
    while len(nan_cases) < 4:
        e = rng.normal(size=9).tolist()
        R = np.array(A.rootsR_X_Y(*e), dtype=complex)
        if np.any(np.isnan(R)):
            nan_cases.append(e)
    out["rootsR_X_Y_nan_inputs"] = nan_cases
    with open(os.path.join(HERE, "autogen_vectors.json"), "w") as fh:
        json.dump(out, fh)
    print({k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
