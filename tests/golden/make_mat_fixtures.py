"""Convert the reference's MATLAB-v5 golden files to JSON fixtures (data only).

Source: legacy/3D_MATLAB/tests/goldenfiles/*.mat of comp-physics/HyQMOM.jl (read with
scipy.io.loadmat, which executes nothing from the file).  Writes tests/golden/mat_goldens.json.
Usage: python tests/golden/make_mat_fixtures.py [/root/reference]
"""
import glob
import json
import os
import sys

import numpy as np
import scipy.io as sio


def tolist(v):
    if hasattr(v, "_fieldnames"):
        return {fn: tolist(getattr(v, fn)) for fn in v._fieldnames}
    a = np.asarray(v)
    if a.dtype.kind in "U":
        return str(a)
    if a.dtype.kind in "iufb":
        a = a.astype(np.float64)
        return a.tolist() if a.ndim else float(a)
    return str(a)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    gdir = os.path.join(ref, "legacy", "3D_MATLAB", "tests", "goldenfiles")
    out = {}
    for f in sorted(glob.glob(os.path.join(gdir, "*.mat"))):
        d = sio.loadmat(f, squeeze_me=True, struct_as_record=False)
        out[os.path.basename(f)[:-4]] = tolist(d["golden_data"].tests)
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "mat_goldens.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", len(out), "golden sets")


if __name__ == "__main__":
    main()
