"""Fixtures from the reference's own golden files for the Nz = 1 runs of run_simulation.

Source: test/goldenfiles/mpi_1rank_{small,medium}.bin of comp-physics/HyQMOM.jl, written by
test/create_golden_files.jl:85-93 (Int64 ranks, Int64 Np, Float64 tmax, Float64 final time,
Int64 steps, then the final state (Np, Np, 35) in column-major order) from
run_simulation(Nx = Ny = Np, tmax, Kn = 1, Ma = 0, flag2D = 0, CFL = 0.5) on one rank
(create_golden_files.jl:33-36: small Np = 20, tmax = 0.1; medium Np = 40, tmax = 0.05, as
the files' headers also say).  The same data
is written to tests/golden/mpi_1rank_<name>.npz: M (Np, Np, 1, 35), tmax, final_time, steps.

    python tests/golden/make_bin_fixtures.py /root/reference
"""
import os
import struct
import sys

import numpy as np


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    for name in ("small", "medium"):
        with open(os.path.join(ref, "test", "goldenfiles", "mpi_1rank_%s.bin" % name), "rb") as f:
            b = f.read()
        nranks, Np = struct.unpack("<qq", b[:16])
        tmax, tfinal = struct.unpack("<dd", b[16:32])
        steps = struct.unpack("<q", b[32:40])[0]
        M = np.frombuffer(b[40:], dtype="<f8")
        assert nranks == 1 and M.size == Np * Np * 35
        M = M.reshape((Np, Np, 1, 35), order="F")
        np.savez_compressed(os.path.join(here, "mpi_1rank_%s.npz" % name), M=M, tmax=tmax,
                            final_time=tfinal, steps=steps, Np=Np)
        print(name, Np, tmax, tfinal, steps)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
