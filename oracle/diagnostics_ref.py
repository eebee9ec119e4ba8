"""CPU reference of the runner's diagnostics (test infrastructure only).

symmetry_maxdiff restates the symmetry check of src/simulation_runner.jl:565-628 on one
host state; vmax_breakdown the first-step vmax breakdown and worst-cell search of
src/simulation_runner.jl:404-461 on the six wave-speed fields.  Plain loops in the
reference's order, for small grids.
"""
import math

import numpy as np


def symmetry_maxdiff(M):
    """M: (Nx, Ny, Nz, 35).  Returns (MaxDiff, per-moment values) as
    simulation_runner.jl:575-628: diagonal cells (i, i, 1) of the first five moments against
    (Nd+1-i, Nd+1-i, 1), Nd = min(Nx, Ny); Diff = a - b for moments 1, 3, 5 and a + b for 2
    and 4; per moment maximum(Diff) / (norm(Diff) + 1); MaxDiff the largest magnitude."""
    Nx, Ny = M.shape[0], M.shape[1]
    nd = min(Nx, Ny)
    diag = np.array([[M[i, i, 0, n] for n in range(5)] for i in range(nd)])
    diff = np.zeros((nd, 5))
    for i in range(nd):
        j = nd - 1 - i
        for n in range(5):
            diff[i, n] = diag[i, n] + diag[j, n] if n % 2 == 1 else diag[i, n] - diag[j, n]
    per = np.zeros(5)
    for n in range(5):
        per[n] = max(diff[:, n]) / (math.sqrt(sum(d * d for d in diff[:, n])) + 1.0)
    return max(abs(per)), per


def vmax_breakdown(vp, Nx, Ny, Nz):
    """vp: (6, N) wave speeds (x min, x max, y min, y max, z min, z max), cell
    c = i + Nx (j + Ny k).  Returns ((vx_max, vy_max, vz_max), (i, j, k) 0-based) as
    simulation_runner.jl:406-409 and the search of :438-461 (loops k, i, j; strict >, so a
    cell whose speeds include NaN is never taken; (-1, -1, -1) when no cell is).  When the
    largest speed is 0 the first cell is returned, where the reference reports none (it
    prints the search only for vmax > 100)."""
    vp = np.asarray(vp, float)
    vdir = tuple(float(np.max(np.abs(vp[2 * d:2 * d + 2]))) for d in range(3))
    worst_v, worst = 0.0, (-1, -1, -1)
    found = False
    for k in range(Nz):
        for i in range(Nx):
            for j in range(Ny):
                c = i + Nx * (j + Ny * k)
                v = np.max(np.abs(vp[:, c]))
                if v > worst_v or (not found and v == worst_v):
                    worst_v, worst, found = v, (i, j, k), True
    return vdir, worst
