"""Shared helpers for the tests (moment orderings, seeded moment generators)."""
import json
import os
from math import comb

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
E35 = [(0,0,0),(1,0,0),(2,0,0),(3,0,0),(4,0,0),(0,1,0),(1,1,0),(2,1,0),(3,1,0),(0,2,0),
       (1,2,0),(2,2,0),(0,3,0),(1,3,0),(0,4,0),(0,0,1),(1,0,1),(2,0,1),(3,0,1),(0,0,2),
       (1,0,2),(2,0,2),(0,0,3),(1,0,3),(0,0,4),(0,1,1),(1,1,1),(2,1,1),(0,2,1),(1,2,1),
       (0,3,1),(0,1,2),(1,1,2),(0,1,3),(0,2,2)]
N28 = ("S300 S400 S110 S210 S310 S120 S220 S030 S130 S040 S101 S201 S301 S102 S202 S003 "
       "S103 S004 S011 S111 S211 S021 S121 S031 S012 S112 S013 S022").split()


def key(n):
    return tuple(int(c) for c in n[1:])


def load_json(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


def gaussian_raw(rho, mean, C, order=4):
    """Exact raw moments of rho * N(mean, C) up to `order` (Isserlis), dict by exponent."""
    def pm(l):
        if not l:
            return 1.0
        a = l[0]
        return sum(C[a][l[t]] * pm(l[1:t] + l[t + 1:]) for t in range(1, len(l)))
    out = {}
    for i in range(order + 1):
        for j in range(order + 1 - i):
            for k in range(order + 1 - i - j):
                idx = [0] * i + [1] * j + [2] * k
                n = len(idx)
                tot = 0.0
                for mask in range(1 << n):
                    sel = [idx[t] for t in range(n) if mask >> t & 1]
                    if len(sel) % 2:
                        continue
                    p = 1.0
                    for t in range(n):
                        if not mask >> t & 1:
                            p *= mean[idx[t]]
                    tot += p * pm(sel)
                out[(i, j, k)] = rho * tot
    return out


def to35(d):
    return np.array([d[e] for e in E35])


def mixture_raw(rng, kmin=2, kmax=4, spread=1.0, lam_lo=0.05, order=4):
    K = int(rng.integers(kmin, kmax + 1))
    w = rng.dirichlet(np.ones(K))
    tot = None
    for c in range(K):
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        lam = np.exp(rng.uniform(np.log(lam_lo), 0, 3))
        C = Q @ np.diag(lam) @ Q.T
        d = gaussian_raw(w[c], rng.uniform(-spread, spread, 3), C, order)
        tot = d if tot is None else {k: tot[k] + d[k] for k in d}
    rho = rng.uniform(0.2, 2.0)
    return {k: rho * v for k, v in tot.items()}
