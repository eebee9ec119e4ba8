"""Host-side initial conditions, mirroring src/initial_conditions.jl of HyQMOM.jl.

CubicRegion, the global cell-centre grid, the strict-overlap cell/cube test and the
region -> Maxwellian moments map.  Produces the (Nx,Ny,Nz,35) state in the reference's
Julia column-major layout (returned as a Fortran-ordered numpy array).
"""
from dataclasses import dataclass
from fractions import Fraction
import math

import numpy as np

from .moments import maxwellian35


@dataclass
class CubicRegion:
    """src/initial_conditions.jl:34-55"""
    center: tuple = (0.0, 0.0, 0.0)
    width: tuple = (math.inf, math.inf, math.inf)
    density: float = 1.0
    velocity: tuple = (0.0, 0.0, 0.0)
    temperature: float = 1.0


def _rat(x):
    """Smallest-denominator rational p/q with float(p/q) == x (continued fractions), as
    Julia's range constructor looks for before falling back to extended precision."""
    f = Fraction(x)
    a = math.floor(f)
    h0, h1, k0, k1 = 1, a, 0, 1
    r = f - a
    for _ in range(64):
        if float(Fraction(h1, k1)) == x:
            return h1, k1
        if r == 0:
            break
        r = 1 / r
        a = math.floor(r)
        r -= a
        h0, h1 = h1, a * h1 + h0
        k0, k1 = k1, a * k1 + k0
    return None


def cell_centers(lo, hi, n):
    """range(lo + d/2, step=d, length=n) with d = (hi-lo)/n, elementwise as Julia builds
    it (src/simulation_runner.jl:192-194): when start and step have exact small rationals,
    each element is the correctly rounded value of (start_n + (i-1) step_n)/den."""
    d = (hi - lo) / n
    a = lo + d / 2
    ra, rs = _rat(a), _rat(d)
    if ra is not None and rs is not None:
        den = ra[1] * rs[1] // math.gcd(ra[1], rs[1])
        sn = ra[0] * (den // ra[1])
        tn = rs[0] * (den // rs[1])
        if abs(den * a) <= 2.0**53 and abs(den * d) <= 2.0**53:
            return np.array([float(Fraction(sn + i * tn, den)) for i in range(n)]), d
    # extended-precision fallback: start + i*step with one rounding
    return np.array([float(Fraction(a) + i * Fraction(d)) for i in range(n)]), d


def cell_overlaps_cube(c, size, cube):
    """src/initial_conditions.jl:281-302 (strict < on each axis)."""
    if any(math.isinf(w) for w in cube.width):
        return True
    for ax in range(3):
        x, dx, cx, wx = c[ax], size[ax], cube.center[ax], cube.width[ax]
        if not (max(x - dx / 2, cx - wx / 2) < min(x + dx / 2, cx + wx / 2)):
            return False
    return True


def region_to_moments(region, r110=0.0, r101=0.0, r011=0.0):
    """src/initial_conditions.jl:224-238"""
    U, V, W = region.velocity
    T = region.temperature
    C = np.array([[T, r110 * math.sqrt(T * T), r101 * math.sqrt(T * T)],
                  [r110 * math.sqrt(T * T), T, r011 * math.sqrt(T * T)],
                  [r101 * math.sqrt(T * T), r011 * math.sqrt(T * T), T]])
    return maxwellian35(region.density, (U, V, W), C)


def initialize_moment_field(Nx, Ny, Nz, background, regions, r110=0.0, r101=0.0, r011=0.0,
                            xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5, zmin=-0.5, zmax=0.5):
    """initialize_moment_field_mpi (src/initial_conditions.jl:152-217) on one domain.
    Returns (M, region_index) with M Fortran-ordered (Nx,Ny,Nz,35)."""
    xm, _ = cell_centers(xmin, xmax, Nx)
    ym, _ = cell_centers(ymin, ymax, Ny)
    zm, _ = cell_centers(zmin, zmax, Nz)
    dx = abs(xm[1] - xm[0]) if Nx > 1 else 0.0
    dy = abs(ym[1] - ym[0]) if Ny > 1 else 0.0
    dz = abs(zm[1] - zm[0]) if Nz > 1 else 0.0
    size = (dx, dy, dz)
    moms = [region_to_moments(background, r110, r101, r011)] + \
           [region_to_moments(r, r110, r101, r011) for r in regions]
    # per-axis overlap masks (the test is separable); later regions overwrite earlier ones
    idx = np.zeros((Nx, Ny, Nz), dtype=np.int32)
    for ri, reg in enumerate(regions, start=1):
        if any(math.isinf(w) for w in reg.width):
            idx[...] = ri
            continue
        mx = np.array([(max(x - dx / 2, reg.center[0] - reg.width[0] / 2) < min(x + dx / 2, reg.center[0] + reg.width[0] / 2)) for x in xm])
        my = np.array([(max(y - dy / 2, reg.center[1] - reg.width[1] / 2) < min(y + dy / 2, reg.center[1] + reg.width[1] / 2)) for y in ym])
        mz = np.array([(max(z - dz / 2, reg.center[2] - reg.width[2] / 2) < min(z + dz / 2, reg.center[2] + reg.width[2] / 2)) for z in zm])
        sel = mx[:, None, None] & my[None, :, None] & mz[None, None, :]
        idx[sel] = ri
    M = np.zeros((Nx, Ny, Nz, 35), order="F")
    for ri, m in enumerate(moms):
        M[idx == ri] = m
    return M, idx


def crossing_jets(Ma, rho_jet=1.0, rho_bg=0.1, T=1.0, side=0.2, offset=0.25):
    """The crossing-jets configuration: background at rest, two cubes of side `side` centred
    at (-offset,-offset,0) and (+offset,+offset,0) moving with +/-(Ma/sqrt2)(1,1,0)
    (the geometry of crossing_jets_ic, src/initial_conditions.jl:377-433)."""
    Uc = Ma / math.sqrt(2.0)
    bg = CubicRegion((0.0, 0.0, 0.0), (math.inf,) * 3, rho_bg, (0.0, 0.0, 0.0), T)
    j1 = CubicRegion((-offset, -offset, 0.0), (side, side, side), rho_jet, (Uc, Uc, 0.0), T)
    j2 = CubicRegion((offset, offset, 0.0), (side, side, side), rho_jet, (-Uc, -Uc, 0.0), T)
    return bg, [j1, j2]
