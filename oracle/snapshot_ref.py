"""CPU reference of the snapshot fields (test infrastructure only): compute_central_field and
compute_standardized_field (src/utils/compute_standardized_field.jl:16-62), M2CS4_35
(src/moments/M2CS4_35.jl) per cell, restated with the binomial expansion of the central
moments in numpy.
"""
import numpy as np

from hyqmom_amd.moments import E35


def central_standardized(M):
    """Per-cell central (C) and standardized (S) moments of a (Nx,Ny,Nz,35) field."""
    from math import comb
    rho = M[..., 0]
    mean = [M[..., 1] / rho, M[..., 5] / rho, M[..., 15] / rho]
    raw = {e: M[..., n] / rho for n, e in enumerate(E35)}
    C = np.empty_like(M)
    for n, (i, j, k) in enumerate(E35):
        s = 0.0
        for p in range(i + 1):
            for q in range(j + 1):
                for r in range(k + 1):
                    s = s + comb(i, p) * comb(j, q) * comb(k, r) * (-mean[0]) ** (i - p) * \
                        (-mean[1]) ** (j - q) * (-mean[2]) ** (k - r) * raw[(p, q, r)]
        C[..., n] = s
    C[..., 0] = 1.0
    C[..., 1] = 0.0; C[..., 5] = 0.0; C[..., 15] = 0.0
    sd = [np.sqrt(np.maximum(C[..., 2], 2.220446049250313e-16)),
          np.sqrt(np.maximum(C[..., 9], 2.220446049250313e-16)),
          np.sqrt(np.maximum(C[..., 19], 2.220446049250313e-16))]
    S = np.empty_like(M)
    for n, (i, j, k) in enumerate(E35):
        S[..., n] = C[..., n] / (sd[0] ** i * sd[1] ** j * sd[2] ** k)
    for n in (2, 9, 19):
        S[..., n] = 1.0
    return C, S


