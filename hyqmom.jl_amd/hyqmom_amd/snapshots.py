"""Snapshot output for run_simulation_with_snapshots.

The reference streams JLD2 (HDF5) groups meta/params, meta/snapshot_interval,
meta/n_snapshots, grid, snapshots/NNNNNN/{M,t,step,S,C} (src/simulation_runner.jl:46-52,
291-318, 676-688).  h5py is not available in this image, so the same keys are written to a
numpy .npz archive ("snapshots/000001/M", ...; np.load reads it).  S and C are the
standardized and central moment fields (compute_standardized_field.jl), computed on the
device by hyq_snapshot_fields.

Like the reference, each snapshot goes to the file when it is taken: the archive is opened
in append mode, the snapshot's arrays are written, and it is closed again, so that the file
is a complete archive after every snapshot (a run that stops early keeps the snapshots it
took) and no snapshot stays in host memory.  meta/n_snapshots and the grid are written by
close().
"""
import zipfile

import numpy as np


class SnapshotWriter:
    def __init__(self, filename, params, interval):
        self.filename = filename
        self.n = 0
        meta = {"meta/snapshot_interval": np.array(interval)}
        for k, v in (params.items() if isinstance(params, dict) else vars(params).items()):
            if isinstance(v, (int, float, bool, str)):
                meta["meta/params/" + k] = np.array(v)
        with zipfile.ZipFile(self.filename, "w", allowZip64=True) as zf:
            self._put(zf, meta)

    @staticmethod
    def _put(zf, arrays):
        for k, v in arrays.items():
            with zf.open(k + ".npy", "w", force_zip64=True) as f:
                np.lib.format.write_array(f, np.asanyarray(v), allow_pickle=False)

    def write(self, M, t, step, C, S):
        """One snapshot: the state M and its central (C) and standardized (S) fields."""
        self.n += 1
        key = "snapshots/%06d/" % self.n
        with zipfile.ZipFile(self.filename, "a", allowZip64=True) as zf:
            self._put(zf, {key + "M": M, key + "t": np.array(t), key + "step": np.array(step),
                           key + "S": S, key + "C": C})

    def close(self, grid):
        arrays = {"meta/n_snapshots": np.array(self.n)}
        for k, v in grid.items():
            arrays["grid/" + k] = np.asarray(v)
        with zipfile.ZipFile(self.filename, "a", allowZip64=True) as zf:
            self._put(zf, arrays)
