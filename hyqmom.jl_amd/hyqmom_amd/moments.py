"""Host-side moment utilities: 35-moment ordering and exact Gaussian (Maxwellian) moments.

maxwellian35 is InitializeM4_35 (src/moments/InitializeM4_35.jl:14-76): the raw moments up
to order 4 of rho * N(mean, C), evaluated in closed form (Isserlis' theorem for the central
moments, binomial shift by the mean)."""
from math import comb

import numpy as np

# src/numerics/Flux_closure35_and_realizable_3D.jl:18-20
E35 = [(0,0,0),(1,0,0),(2,0,0),(3,0,0),(4,0,0),(0,1,0),(1,1,0),(2,1,0),(3,1,0),(0,2,0),
       (1,2,0),(2,2,0),(0,3,0),(1,3,0),(0,4,0),(0,0,1),(1,0,1),(2,0,1),(3,0,1),(0,0,2),
       (1,0,2),(2,0,2),(0,0,3),(1,0,3),(0,0,4),(0,1,1),(1,1,1),(2,1,1),(0,2,1),(1,2,1),
       (0,3,1),(0,1,2),(1,1,2),(0,1,3),(0,2,2)]
NAMES35 = ["M%d%d%d" % e for e in E35]


def _gauss_central(C, e):
    idx = [0] * e[0] + [1] * e[1] + [2] * e[2]
    if len(idx) % 2:
        return 0.0

    def pm(l):
        if not l:
            return 1.0
        a = l[0]
        return sum(C[a][l[t]] * pm(l[1:t] + l[t + 1:]) for t in range(1, len(l)))
    return pm(idx)


def maxwellian35(rho, mean, C):
    C = np.asarray(C, dtype=float)
    cen = {}
    for i in range(5):
        for j in range(5 - i):
            for k in range(5 - i - j):
                cen[(i, j, k)] = _gauss_central(C, (i, j, k))
    u, v, w = mean
    out = np.zeros(35)
    for n, (i, j, k) in enumerate(E35):
        s = 0.0
        for p in range(i + 1):
            for q in range(j + 1):
                for r in range(k + 1):
                    s += comb(i, p) * comb(j, q) * comb(k, r) * u ** (i - p) * v ** (j - q) * \
                         w ** (k - r) * cen[(p, q, r)]
        out[n] = rho * s
    return out
