"""Snapshot output for run_simulation_with_snapshots.

The reference streams JLD2 (HDF5) groups meta/params, meta/snapshot_interval,
meta/n_snapshots, grid, snapshots/NNNNNN/{M,t,step,S,C} (src/simulation_runner.jl:46-52,
291-318, 676-688).  h5py is not available in this image, so the same keys are written to a
numpy .npz archive ("snapshots/000001/M", ...).  S and C are the standardized and central
moment fields (compute_standardized_field.jl), computed on the device by
hyq_snapshot_fields.
"""
import numpy as np


class SnapshotWriter:
    def __init__(self, filename, params, interval):
        self.filename = filename
        self.data = {"meta/snapshot_interval": np.array(interval)}
        for k, v in (params.items() if isinstance(params, dict) else vars(params).items()):
            if isinstance(v, (int, float, bool, str)):
                self.data["meta/params/" + k] = np.array(v)
        self.n = 0

    def write(self, M, t, step, C, S):
        """One snapshot: the state M and its central (C) and standardized (S) fields."""
        self.n += 1
        key = "snapshots/%06d/" % self.n
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
