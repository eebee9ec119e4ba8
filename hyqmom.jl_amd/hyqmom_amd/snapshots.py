"""Snapshot output for run_simulation_with_snapshots.

The reference streams JLD2 (HDF5) groups meta/params, meta/snapshot_interval,
meta/n_snapshots, grid, snapshots/NNNNNN/{M,t,step,S,C} (src/simulation_runner.jl:46-52,
291-318, 676-688).  h5py is not available in this image, so the same keys are written to a
numpy .npz archive ("snapshots/000001/M", ...).  S and C are the standardized and central
moment fields (compute_standardized_field.jl).
"""
import numpy as np


def central_standardized(M):
    """Per-cell central (C) and standardized (S) moments of a (Nx,Ny,Nz,35) field."""
    from .moments import E35
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


class SnapshotWriter:
    def __init__(self, filename, params, interval):
        self.filename = filename
        self.data = {"meta/snapshot_interval": np.array(interval)}
        for k, v in (params.items() if isinstance(params, dict) else vars(params).items()):
            if isinstance(v, (int, float, bool, str)):
                self.data["meta/params/" + k] = np.array(v)
        self.n = 0

    def write(self, M, t, step):
        self.n += 1
        key = "snapshots/%06d/" % self.n
        C, S = central_standardized(M)
        self.data[key + "M"] = M
        self.data[key + "t"] = np.array(t)
        self.data[key + "step"] = np.array(step)
        self.data[key + "S"] = S
        self.data[key + "C"] = C

    def close(self, grid):
        self.data["meta/n_snapshots"] = np.array(self.n)
        for k, v in grid.items():
            self.data["grid/" + k] = np.asarray(v)
        np.savez(self.filename, **self.data)
