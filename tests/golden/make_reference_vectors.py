"""Write tests/golden/reference_vectors.npz: per-cell input/output vectors of the reference's
own per-cell functions, evaluated through the Python translation of its Julia sources
(oracle/_ref/reference_jl.py, made by `python oracle/ref_julia.py /root/reference`).

The cells are chosen so that every branch of the closure, realizability projection and wave
speed path is exercised, the oracle's branch flags naming the branch (univariate clamp,
edge/corner, faces, flag220, 2D roots, 4x4-minor roots, x/y/z eigenvalue corrections,
z negative-variance re-projection, density clamp), with positivity on and off.  Each vector
also stores the reference's own sensitivity: the largest change of its outputs over twelve
random relative perturbations of the input of up to 1 and 4 ulps (relative to the output's
scale), so that a test can tell a rounding-level knife edge from a real difference.

Run in the build container (the reference is read here; the GPU box never sees it):
    python oracle/ref_julia.py /root/reference
    python tests/golden/make_reference_vectors.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "oracle", "_ref"),
                os.path.join(ROOT, "hyqmom.jl_amd"), os.path.dirname(HERE)]

import oracle as O  # noqa: E402
import reference_jl as R  # noqa: E402
from julia_rt import JArr, JuliaError, POSITIVITY_ENABLED  # noqa: E402
from _common import E35, gaussian_raw, to35  # noqa: E402
from hyqmom_amd import workloads as W  # noqa: E402

F = dict(UNIV=1 << 0, EDGE=1 << 1, FACES=1 << 2, F220=1 << 3, R2D=1 << 4, R4=1 << 5,
         ECX=1 << 6, ECY=1 << 7, ECZ=1 << 8, ZNEG=1 << 9, NONFIN=1 << 10, RHO=1 << 11)
PER_BIT = 16
PLAIN = 24


# ---------------------------------------------------------------------------------------
# candidate cells
def points_raw(w, pts):
    """Raw moments (order <= 4) of a weighted set of velocity points."""
    pts = np.asarray(pts, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    return np.array([np.sum(w * pts[:, 0] ** i * pts[:, 1] ** j * pts[:, 2] ** k) for (i, j, k) in E35])


def degenerate_cells(rng, n):
    """Distributions on a line or plane of velocity space: |S110|, |S101| or |S011| at 1 up
    to rounding (edges and corners), and two-point marginals (H at its floor)."""
    out = []
    for t in range(n):
        kind = t % 5
        m = int(rng.integers(2, 6))
        s = rng.normal(size=m)
        w = rng.dirichlet(np.ones(m)) * rng.uniform(0.3, 2.0)
        c = rng.uniform(-1, 1, 3)
        sg = rng.choice([-1.0, 1.0], 3)
        if kind == 0:    # u = +-v: xy edge
            pts = np.stack([c[0] + s, c[1] + sg[0] * s, c[2] + rng.normal(size=m)], 1)
        elif kind == 1:  # u = +-w: xz edge
            pts = np.stack([c[0] + s, c[1] + rng.normal(size=m), c[2] + sg[1] * s], 1)
        elif kind == 2:  # v = +-w: yz edge
            pts = np.stack([c[0] + rng.normal(size=m), c[1] + s, c[2] + sg[2] * s], 1)
        elif kind == 3:  # a line in 3D: corner
            pts = np.stack([c[0] + s, c[1] + sg[0] * s, c[2] + sg[1] * s], 1)
        else:            # two-point marginals, independent axes
            a = rng.choice([-1.0, 1.0], (m, 3)) * rng.uniform(0.2, 1.5, 3)
            pts = c + a
        out.append(points_raw(w, pts))
    return np.array(out)


def edge_cells(rng, base, n):
    """Advected-like states whose correlation coefficients leave [-1, 1] by a clear margin
    (1e-6 to 1e-2): one pair (an edge of the edge/corner correction) or all three (a
    corner), so that the branch does not hinge on rounding."""
    out = []
    pairs = [(6, 1, 5, 2, 9), (16, 1, 15, 2, 19), (25, 5, 15, 9, 19)]  # M110, M101, M011
    for t in range(n):
        m = base[rng.integers(len(base))].copy()
        for (ix, a, b, va, vb) in (pairs if t % 4 == 3 else [pairs[t % 4 % 3]]):
            r0 = m[0]
            ua, ub = m[a] / r0, m[b] / r0
            ca, cb = m[va] / r0 - ua * ua, m[vb] / r0 - ub * ub
            c = rng.choice([-1.0, 1.0]) * (1 + np.exp(rng.uniform(np.log(1e-6), np.log(1e-2)))) * np.sqrt(ca * cb)
            m[ix] = r0 * (c + ua * ub)
        out.append(m)
    return np.array(out)


def skewed_cells(rng, n):
    """A far, light component: |S3| beyond s3max = 4 + |Ma|/2 (univariate clamp)."""
    out = []
    for _ in range(n):
        C = np.eye(3) * rng.uniform(0.05, 0.5)
        far = rng.normal(size=3)
        far *= rng.uniform(8, 30) / np.linalg.norm(far)
        a = gaussian_raw(1.0 - 1e-3, (0, 0, 0), C, 4)
        b = gaussian_raw(1e-3 * rng.uniform(0.2, 1.0), tuple(far), np.eye(3) * 0.01, 4)
        out.append(to35({k: a[k] + b[k] for k in a}))
    return np.array(out)


def jets_states(N=(24, 24, 12), Ma=1.5, Kn=0.01, steps=(0, 3, 6, 8)):
    """Cells of the crossing-jets state (oracle trajectory) at a few steps."""
    import hyqmom_amd as H
    bg, jets = H.crossing_jets(Ma)
    p = dict(Nx=N[0], Ny=N[1], Nz=N[2], tmax=10.0, Kn=Kn, Ma=Ma, flag2D=0, CFL=0.7, Nmom=35,
             nnmax=100, dtmax=1e-2, rhol=1.0, rhor=0.1, T=1.0, r110=0.0, r101=0.0, r011=0.0,
             symmetry_check_interval=0, homogeneous_z=False, debug_output=False,
             use_custom_ic=True, ic_background=bg, ic_jets=jets)
    A = np.asfortranarray(H.initial_state(p))
    out, t, done = [], 0.0, 0
    for s in steps:
        if s > done:
            A, t, _, _, rc = O.run(A, *N, Kn, Ma, 0, 0.7, 1e-2, 1 / N[0], 1 / N[1], 1 / N[2], True,
                                   10.0, s - done, t0=t, nn0=0)
            assert rc == 0
            done = s
        out.append(A.reshape(-1, 35, order="F").copy())
    return np.concatenate(out)


def negvar_cells(rng, base, n):
    """Advected-like states with a negative variance estimate in one direction (minus 1% to
    50% of the cell's variance)."""
    out = []
    for t in range(n):
        m = base[rng.integers(len(base))].copy()
        ax = [2, 9, 19][t % 3]
        mean = [1, 5, 15][t % 3]
        var = m[ax] / m[0] - (m[mean] / m[0]) ** 2
        m[ax] = m[0] * ((m[mean] / m[0]) ** 2 - var * rng.uniform(0.01, 0.5))
        out.append(m)
    return np.array(out)


# ---------------------------------------------------------------------------------------
def _flat(res):
    return np.concatenate([np.atleast_1d(np.asarray(x.a if isinstance(x, JArr) else x, dtype=np.float64)).ravel()
                           for x in res])


def evaluate(fn, x, extra):
    try:
        return _flat(fn(JArr(np.array(x, dtype=np.float64)), *extra))
    except JuliaError:
        return None


NEAR = 12  # perturbed evaluations per vector


def sensitivity(fn, x, extra, y, rng):
    """(largest relative change of the outputs, the perturbed outputs [NEAR][len(y)])."""
    sc = max(np.abs(y).max(), 1e-300)
    worst, near = 0.0, np.full((NEAR, len(y)), np.nan)
    for t in range(NEAR):
        ulps = 1.0 if t < NEAR // 2 else 4.0
        xp = x * (1 + rng.uniform(-1.0, 1.0, size=x.shape) * ulps * 2.0 ** -52)
        yp = evaluate(fn, xp, extra)
        if yp is None or yp.shape != y.shape or not np.all(np.isfinite(yp)):
            worst = np.inf
            continue
        near[t] = yp
        worst = max(worst, np.abs(yp - y).max() / sc)
    return worst, near


def select(cands, flagfn):
    """Indices: up to PER_BIT cells per flag bit, plus PLAIN cells with no flag."""
    fl = np.array([flagfn(c) for c in cands], dtype=np.uint32)
    take = []
    for b in F.values():
        take += list(np.nonzero(fl & b)[0][:PER_BIT])
    take += list(np.nonzero(fl == 0)[0][:PLAIN])
    return sorted(set(take)), fl


def build_group(name, cands, fn, extra, flagfn, rng, log):
    idx, fl = select(cands, flagfn)
    X, Y, S, FL, NR = [], [], [], [], []
    for i in idx:
        y = evaluate(fn, cands[i], extra)
        if y is None or not np.all(np.isfinite(y)):
            continue
        sens, near = sensitivity(fn, cands[i], extra, y, rng)
        X.append(cands[i]); Y.append(y); FL.append(fl[i]); S.append(sens)
        # the perturbed outputs are kept for knife-edge vectors only (NaN otherwise)
        NR.append(near if sens > 1e-9 else np.full_like(near, np.nan))
    X, Y, S, FL, NR = map(np.array, (X, Y, S, FL, NR))
    hist = {k: int(((FL & b) != 0).sum()) for k, b in F.items() if ((FL & b) != 0).any()}
    log.append("%s: %d vectors (%d knife-edge), flags %s" % (name, len(X), int((S > 1e-9).sum()), hist))
    return {name + "_x": X, name + "_y": Y, name + "_sens": S, name + "_flags": FL, name + "_near": NR}


def main():
    t0 = time.time()
    rng = np.random.default_rng(20260531)
    base = jets_states()
    stress = np.asarray(W.stress_batch(6000, seed=11))
    cands = np.concatenate([stress, base[rng.permutation(len(base))[:6000]],
                            edge_cells(rng, base, 400), degenerate_cells(rng, 200),
                            skewed_cells(rng, 60)])
    log, out = [], {}

    # Flux_closure35_and_realizable_3D (3 x 35 fluxes + projected M), flag2D 0 and 1
    cl = lambda f2, Ma: (lambda c: O.flux_closure35(c, f2, Ma)[4])
    out.update(build_group("closure", cands, R.Flux_closure35_and_realizable_3D, (0, 1.5),
                           cl(0, 1.5), rng, log))
    out.update(build_group("closure2d", cands[::7], R.Flux_closure35_and_realizable_3D, (1, 1.0),
                           cl(1, 1.0), rng, log))

    # eigenvalues6_hyperbolic_3D, x (axis 1) and y (axis 2): (v6min, v6max, Mr)
    for ax in (1, 2):
        ef = lambda c, ax=ax: O.eigenvalues6(c, ax, 0, 1.5)[3]
        out.update(build_group("eig%d" % ax, cands, R.eigenvalues6_hyperbolic_3D, (ax, 0, 1.5),
                               ef, rng, log))

    # eigenvalues6z_hyperbolic_3D with positivity on / off, including negative-variance
    # and non-positive density inputs
    zc = np.concatenate([cands[::3], negvar_cells(rng, base, 300)])
    rho = base[rng.integers(len(base), size=40)].copy()
    rho[:, 0] = -rho[:, 0] * rng.uniform(0, 1e-3, 40)
    zc = np.concatenate([zc, rho])
    for pos in (True, False):
        POSITIVITY_ENABLED.value = pos
        zf = lambda c, pos=pos: O.eigenvalues6z(c, 0, 1.5, pos)[3]
        out.update(build_group("eigz_pos%d" % pos, zc, R.eigenvalues6z_hyperbolic_3D, (0, 1.5),
                               zf, rng, log))
    POSITIVITY_ENABLED.value = True

    # realizable_3D directly, with faces inputs (S2 at zero after the 0.9999 shrink)
    C, S = zip(*(O.m2cs4(c) for c in cands[::5]))
    names = ("S300 S400 S110 S210 S310 S120 S220 S030 S130 S040 S101 S201 S301 S102 S202 S003 "
             "S103 S004 S011 S111 S211 S021 S121 S031 S012 S112 S013 S022").split()
    pos35 = {e: n for n, e in enumerate(E35)}
    idx28 = [pos35[tuple(int(ch) for ch in nm[1:])] for nm in names]
    s28 = np.array([np.asarray(s)[idx28] for s in S])
    faces = s28[rng.integers(len(s28), size=60)].copy()
    for t, row in enumerate(faces):
        sgn = rng.choice([-1.0, 1.0], 2)
        r = 1.0 / 0.9999 * (1 - rng.uniform(0, 2e-7))
        row[2], row[10], row[18] = r, sgn[0] * r, sgn[0] * sgn[1] * r * (1 if t % 2 else 1)
    s28 = np.concatenate([s28, faces])
    r3 = lambda v: O.realizable_3D_flags(v)[2]
    out.update(build_group("real3d", s28, lambda v: R.realizable_3D(*list(v)), (), r3, rng, log))

    # check2D (collapse or realizable_2D with the L3 < 0 roots branch) on random planes
    p10 = []
    for c in cands[::4]:
        _, s = O.m2cs4(c)
        p10.append([s[3], s[4], s[6], s[7], s[8], s[10], s[11], s[12], s[13], s[14]])
    p10 = np.array(p10)
    c2 = lambda v: (F["R2D"] if _l3(v) < 0 else 0)
    out.update(build_group("check2d", p10, lambda v: R.check2D(*list(v)), (), c2, rng, log))

    # collision35 (positivity on / off) and the 1D closure and Moments5_3D
    col = cands[rng.permutation(len(cands))[:80]]
    for pos in (True, False):
        POSITIVITY_ENABLED.value = pos
        out.update(build_group("bgk_pos%d" % pos, col, R.collision35, (3.1e-3, 0.01),
                               lambda c: 0, rng, log))
    POSITIVITY_ENABLED.value = True
    m5 = np.array([c[[0, 1, 2, 3, 4]] for c in cands[rng.permutation(len(cands))[:60]]])
    out.update(build_group("closure1d", m5, R.closure_and_eigenvalues, (), lambda c: 0, rng, log))
    out.update(build_group("moments5", cands[rng.permutation(len(cands))[:80]], R.Moments5_3D, (),
                           lambda c: 0, rng, log))

    path = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(path, **out)
    for ln in log:
        print(ln)
    print("wrote %s (%.0f s)" % (path, time.time() - t0))


def _l3(v):
    S30, S40, S11, S21, S31, S12, S22, S03, S13, S04 = v
    return O.delta2starchol_L3([S03, S04, S11, S12, S13, S21, S22, S30, S31, S40])


if __name__ == "__main__":
    main()
