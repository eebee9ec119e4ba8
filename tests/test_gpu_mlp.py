"""The warm-start network on the GPU (bf16 MFMA kernels) against its CPU reference, at both
hidden widths (32, the shipped network, and 128, the width of the warm-start configuration),
the training dataset against the oracle's eigen solve, and the mixed_learned / mixed_newton
precision modes against the plain mixed step."""
import numpy as np
import pytest

import mlp_ref as R
import oracle as O

pytestmark = pytest.mark.gpu


def _dev(a, dtype=np.float32):
    from gpu_util import to_dev
    return to_dev(np.ascontiguousarray(a, dtype=dtype))


WIDTHS = [32, 128]


@pytest.mark.parametrize("hidden", WIDTHS)
@pytest.mark.parametrize("n", [1, 37, 1000])
def test_forward_matches_reference(n, hidden):
    """Random weights: any mistake in the MFMA operand/result lane layout or the strides
    gives O(1) errors; bf16 rounding of the operands is reproduced by the reference, so
    the remaining difference is FP32 accumulation order (and rare bf16 rounding flips)."""
    from hyqmom_amd.warmstart import WarmStartMLP
    blob = R.random_blob(11, hidden=hidden)
    X = np.random.default_rng(12).standard_normal((n, 16)).astype(np.float32)
    m = WarmStartMLP(blob)
    assert m.hidden == hidden
    P = m.infer(_dev(X)).cpu().numpy()
    ref = R.predict(blob, X, use_bf16=True)
    ysd = R.unpack(blob)["ysd"]
    err = np.abs(P - ref) / ysd
    print("H=%d n=%d forward max err %.3e mean %.3e" % (hidden, n, err.max(), err.mean()))
    assert err.max() < 2e-2 and err.mean() < 2e-3
    # and the reference differs from the unrounded FP64 network by more than that
    ref64 = R.predict(blob, X, use_bf16=False)
    assert np.abs(P - ref64).max() / ysd.max() > 0 or n == 1


@pytest.mark.parametrize("hidden", WIDTHS)
@pytest.mark.parametrize("n", [1, 63, 4099, 70001])
def test_fused_inference_equals_layers(n, hidden):
    """The fused kernel (weights packed in LDS, activations per wave) rounds and accumulates
    exactly like the layer-by-layer kernels: bitwise equal predictions (n = 70001 also runs
    the launch-shape timing, which keeps the faster of two equal-valued shapes)."""
    import torch
    from hyqmom_amd.warmstart import WarmStartMLP
    blob = R.random_blob(13, hidden=hidden)
    X = np.random.default_rng(14).standard_normal((n, 16)).astype(np.float32)
    m = WarmStartMLP(blob)
    Xd = _dev(X)
    a = m.infer(Xd).cpu().numpy()
    b = m.infer_layers(Xd).cpu().numpy()
    assert np.array_equal(a, b), np.abs(a - b).max()
    # a parameter change reaches the packed copy
    blob2 = R.random_blob(15, hidden=hidden)
    m.set_blob(blob2)
    c = m.infer(Xd).cpu().numpy()
    assert np.array_equal(c, m.infer_layers(Xd).cpu().numpy())
    assert not np.array_equal(a, c)


@pytest.mark.parametrize("hidden", WIDTHS)
@pytest.mark.parametrize("shape", [0, 1, 2])
def test_every_fused_shape_equals_layers(shape, hidden):
    """Each launch shape of the fused inference gives the layer-by-layer predictions bit for
    bit, masked rows aside:
    with the mask a 16-row tile whose rows all have column 15 = 0 gets NaN."""
    from hyqmom_amd import _lib
    from hyqmom_amd.warmstart import WarmStartMLP
    n = 5000
    X = np.random.default_rng(21).standard_normal((n, 16)).astype(np.float32)
    X[:, 15] = 1.0
    X[1600:1632, 15] = 0.0  # two whole 16-row tiles to skip
    m = WarmStartMLP(R.random_blob(22, hidden=hidden))
    m.set_shape(shape)
    Xd = _dev(X)
    want = m.infer_layers(Xd).cpu().numpy()
    got = m.infer(Xd).cpu().numpy()
    assert np.array_equal(got, want), (shape, np.abs(got - want).max())
    import ctypes
    import torch
    P = torch.empty((n, 2), dtype=torch.float32, device="cuda")
    # the step's masked inference (exported for the step code, not in the public header)
    fn = _lib.lib().hyq_mlp_infer_masked_async
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    fn.restype = ctypes.c_int
    rc = fn(m.handle, Xd.data_ptr(), n, P.data_ptr(), None)
    assert rc == 0
    torch.cuda.synchronize()
    Pm = P.cpu().numpy()
    skip = np.zeros(n, bool)
    skip[1600:1632] = True
    assert np.isnan(Pm[skip]).all()
    assert np.array_equal(Pm[~skip], want[~skip])


@pytest.mark.parametrize("hidden", WIDTHS)
def test_gradient_matches_reference(hidden):
    from hyqmom_amd.warmstart import WarmStartMLP
    blob = R.random_blob(21, hidden=hidden)
    rng = np.random.default_rng(22)
    n = 777
    X = rng.standard_normal((n, 16)).astype(np.float32)
    T = rng.standard_normal((n, 2)).astype(np.float32) * 0.5 + np.array([-2.0, 2.0], np.float32)
    m = WarmStartMLP(blob)
    loss, g = m.grad(_dev(X), _dev(T))
    lref, gref = R.loss_and_grad(blob, X, T, use_bf16=True)
    assert abs(loss - lref) <= 1e-3 * lref, (loss, lref)
    o = 0
    for name, shp in R.shapes(hidden):
        k = int(np.prod(shp))
        a, b = g[o:o + k], gref[o:o + k]
        o += k
        nb = np.linalg.norm(b)
        if nb == 0:
            assert np.all(a == 0), name
            continue
        rel = np.linalg.norm(a - b) / nb
        print("%s grad rel err %.3e" % (name, rel))
        assert rel < 1e-2, (name, rel)


@pytest.mark.parametrize("hidden", WIDTHS)
def test_training_reduces_loss(hidden):
    """Adam on a smooth synthetic regression: the loss falls by an order of magnitude."""
    import torch
    from hyqmom_amd import warmstart as W
    rng = np.random.default_rng(31)
    n = 20000
    X = rng.standard_normal((n, 16)).astype(np.float32)
    X[:, 15] = 0.0
    T = np.stack([-1.5 - 0.3 * np.tanh(X[:, 0] * X[:, 3]) - 0.2 * X[:, 4] ** 2,
                  1.5 + 0.3 * np.sin(X[:, 6]) + 0.2 * X[:, 7] * X[:, 8]], axis=1).astype(np.float32)
    Xd, Td = _dev(X), _dev(T)
    m, hist = W.train(Xd, Td, steps=400, batch=2048, lr=2e-3, lr_final=2e-4, seed=1, hidden=hidden)
    print("H=%d loss history" % hidden, hist[0], hist[-1])
    assert hist[-1][1] < 0.1 * hist[0][1]
    b1 = m.blob()
    m2 = W.WarmStartMLP(b1)
    assert np.array_equal(m2.blob(), b1)
    P1 = m.infer(Xd[:100]).cpu().numpy()
    P2 = m2.infer(Xd[:100]).cpu().numpy()
    assert np.array_equal(P1, P2)


def test_plane_dataset_labels_match_oracle():
    """Features are the 15 standardized plane moments in jacobian6's argument order; labels
    are the extreme eigenvalues of that Jacobian (oracle: LAPACK-family QR)."""
    from hyqmom_amd.warmstart import plane_dataset
    from _common import mixture_raw, to35
    rng = np.random.default_rng(41)
    M = np.array([to35(mixture_raw(rng, spread=1.5, lam_lo=0.05)) for _ in range(200)])
    X, L, V = plane_dataset(_dev(M.T, np.float64))
    X, L, V = X.cpu().numpy(), L.cpu().numpy(), V.cpu().numpy()
    n = len(M)
    assert X.shape == (5 * n, 16) and np.all(X[:, 15] == 0)
    assert np.all(X[:, [0, 2, 9]] == 1) and np.all(X[:, [1, 5]] == 0)
    checked = 0
    for s in range(5 * n):
        ev = O.eigvals(O.jacobian6(X[s, :15].astype(np.float64)))
        real = np.max(np.abs(ev.imag)) <= 1e-6 and np.all(np.isfinite(ev))
        if V[s]:
            assert real
            r = np.sort(ev.real)
            assert abs(L[s, 0] - r[0]) <= 1e-4 * (1 + abs(r[0])), (s, L[s], r)
            assert abs(L[s, 1] - r[-1]) <= 1e-4 * (1 + abs(r[-1])), (s, L[s], r)
            checked += 1
        elif np.max(np.abs(ev.imag)) > 1e-3:
            pass  # complex: invalid on both sides
    assert checked > 0.5 * 5 * n


def _params(N, Ma, Kn, prec):
    import hyqmom_amd as H
    bg, jets = H.crossing_jets(Ma)
    return dict(Nx=N[0], Ny=N[1], Nz=N[2], tmax=10.0, Kn=Kn, Ma=Ma, flag2D=0, CFL=0.7, Nmom=35,
                nnmax=100, dtmax=1e-2, rhol=1.0, rhor=0.1, T=1.0, r110=0.0, r101=0.0, r011=0.0,
                symmetry_check_interval=0, homogeneous_z=False, debug_output=False,
                use_custom_ic=True, ic_background=bg, ic_jets=jets, precision=prec)


def test_newton_and_learned_modes_match_mixed():
    """From one state, mixed_newton and mixed_learned (an untrained network: any warm start
    is only a Laguerre seed) give the wave speeds of the full solve to the iteration tolerance
    (so dt agrees to 1e-12), hence the same step as the mixed mode up to FP32 rounding."""
    import hyqmom_amd as H
    from hyqmom_amd.warmstart import WarmStartMLP
    N = (24, 24, 12)
    p = _params(N, 1.5, 0.01, "mixed")
    M0 = np.asfortranarray(H.initial_state(p))
    out = {}
    for prec in ("mixed", "mixed_newton", "mixed_learned"):
        q = dict(p, precision=prec)
        if prec == "mixed_learned":
            q["warmstart_model"] = WarmStartMLP(R.random_blob(51))
        s = H.Solver(q)
        s.set_state(M0)
        dt = s.step()
        out[prec] = (s.get_state(), s.wave_speeds(), dt, s.flags())
        s.close()
    Mm, vm, dtm, _ = out["mixed"]
    for prec in ("mixed_newton", "mixed_learned"):
        Mx, vx, dtx, fx = out[prec]
        assert abs(dtx - dtm) <= 1e-12 * dtm, (prec, dtx, dtm)
        ev = np.abs(vx - vm) / np.maximum(np.abs(vm).max(axis=1, keepdims=True), 1e-300)
        print(prec, "wave-speed max rel diff %.2e, fallback cells %d" % (ev.max(), ((fx >> 12) & 1).sum()))
        assert ev.max() < 2e-7  # mixed modes keep the wave speeds in FP32
        scale = np.abs(Mm).max(axis=(0, 1, 2))
        e = (np.abs(Mx - Mm) / scale).max(axis=3)
        assert e.max() < 1e-6, (prec, e.max())


def test_trained_network_reduces_iterations():
    """What the trained network (hyqmom.jl_amd/models/warmstart_mlp.f32) gains: on a
    developed crossing-jets state, the Laguerre iterations of the extreme wave speeds with
    the network's warm starts against the analytic bounds, at identical results."""
    import os
    import hyqmom_amd as H
    from hyqmom_amd import _lib
    from hyqmom_amd.warmstart import DEFAULT_MODEL, WarmStartMLP
    assert os.path.exists(DEFAULT_MODEL)
    N = (60, 60, 30)
    p = _params(N, 1.5, 0.01, "fp64")
    s = H.Solver(p)
    s.set_state(H.initial_state(p))
    s.run(10.0, 25)
    M = s.get_state()
    s.close()
    out = {}
    for prec in ("mixed_newton", "mixed_learned"):
        q = dict(p, precision=prec)
        if prec == "mixed_learned":
            q["warmstart_model"] = WarmStartMLP.load(DEFAULT_MODEL)
        s = H.Solver(q)
        s.set_state(M)
        s.step()
        fl = s.flags()
        out[prec] = (_lib.newton_iterations(fl).mean(), ((fl >> 12) & 1).mean(), s.wave_speeds(), s.get_state())
        s.close()
    it_n, fb_n, v_n, M_n = out["mixed_newton"]
    it_l, fb_l, v_l, M_l = out["mixed_learned"]
    print("iterations per cell: analytic start %.2f, network start %.2f (%.0f%% fewer); fallback "
          "cells %.4f / %.4f" % (it_n, it_l, 100 * (1 - it_l / it_n), fb_n, fb_l))
    assert it_l < 0.85 * it_n
    assert fb_l == fb_n  # the hyperbolicity decision does not depend on the start
    scale = np.abs(v_n).max(axis=1, keepdims=True)
    assert (np.abs(v_l - v_n) / scale).max() < 2e-7
    sc = np.abs(M_n).max(axis=(0, 1, 2))
    assert (np.abs(M_l - M_n) / sc).max() < 1e-6


@pytest.mark.parametrize("hidden", WIDTHS)
@pytest.mark.parametrize("shape", [0, 1, 2])
def test_bf16_rows_give_the_same_predictions(shape, hidden):
    """The step's projection kernel hands the network its input rows already normalized and
    rounded to bf16 (the same FP32 subtraction and division, then round to nearest even);
    the fused kernel fed those rows gives the float rows' predictions bit for bit, masked
    tiles aside.  The rows here are made on the host with numpy's IEEE FP32 arithmetic."""
    import ctypes
    import torch
    from hyqmom_amd import _lib
    from hyqmom_amd.warmstart import WarmStartMLP
    n = 5000
    X = np.random.default_rng(23).standard_normal((n, 16)).astype(np.float32)
    X[:, 15] = 1.0
    X[1600:1632, 15] = 0.0
    blob = R.random_blob(24, hidden=hidden)
    m = WarmStartMLP(blob)
    m.set_shape(shape)
    npar = R.nparam(hidden)
    mu = blob[npar:npar + 16].astype(np.float32)
    sd = blob[npar + 16:npar + 32].astype(np.float32)
    Xn = X.copy()
    Xn[:, :15] = (X[:, :15] - mu[:15]) / sd[:15]
    bits = R.bf16(Xn).view(np.uint32) >> 16
    Xb = torch.from_numpy(bits.astype(np.uint16).view(np.int16).copy()).cuda()
    Xd = torch.from_numpy(X).cuda()
    Pf = torch.empty((n, 2), dtype=torch.float32, device="cuda")
    Pb = torch.empty((n, 2), dtype=torch.float32, device="cuda")
    L = _lib.lib()
    for name, src, dst in (("hyq_mlp_infer_masked_async", Xd, Pf),
                           ("hyq_mlp_infer_bf16_masked_async", Xb, Pb)):
        fn = getattr(L, name)
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        fn.restype = ctypes.c_int
        assert fn(m.handle, src.data_ptr(), n, dst.data_ptr(), None) == 0
    torch.cuda.synchronize()
    a, b = Pf.cpu().numpy(), Pb.cpu().numpy()
    skip = np.zeros(n, bool)
    skip[1600:1632] = True
    assert np.isnan(b[skip]).all()
    assert np.array_equal(a[~skip], b[~skip]), np.abs(a - b)[~skip].max()


@pytest.mark.parametrize("hidden", WIDTHS)
def test_segmented_rows_give_the_same_predictions(hidden):
    """The step hands the network only the planes that need it: each wave of the projection
    kernel fills the front of its own segment of 320 rows and writes its count
    (hyq_mlp_infer_bf16_segments_async reads the counts on the device).  Each listed row's
    prediction lands at its plane index, bit for bit the prediction of the unsegmented
    kernel for the same row; rows past a segment's count and unlisted planes are not
    written.  Counts 0, partial and full are covered."""
    import ctypes
    import torch
    from hyqmom_amd import _lib
    from hyqmom_amd.warmstart import WarmStartMLP
    nseg, srows, nplanes = 17, 320, 20000
    n = nseg * srows
    rng = np.random.default_rng(29)
    X = rng.standard_normal((n, 16)).astype(np.float32)
    X[:, 15] = 1.0
    counts = rng.integers(0, srows + 1, nseg).astype(np.uint32)
    counts[0], counts[1], counts[2] = 0, srows, 1
    blob = R.random_blob(25, hidden=hidden)
    m = WarmStartMLP(blob)
    npar = R.nparam(hidden)
    mu = blob[npar:npar + 16].astype(np.float32)
    sd = blob[npar + 16:npar + 32].astype(np.float32)
    Xn = X.copy()
    Xn[:, :15] = (X[:, :15] - mu[:15]) / sd[:15]
    bits = (R.bf16(Xn).view(np.uint32) >> 16).astype(np.uint16).view(np.int16)
    Xb = torch.from_numpy(bits.copy()).cuda()
    idx = rng.choice(nplanes, n, replace=False).astype(np.uint32)
    Id = torch.from_numpy(idx.view(np.int32).copy()).cuda()
    Sd = torch.from_numpy(counts.view(np.int32).copy()).cuda()
    Pref = torch.empty((n, 2), dtype=torch.float32, device="cuda")
    Pc = torch.full((nplanes, 2), 7.0, dtype=torch.float32, device="cuda")
    L = _lib.lib()
    fm = L.hyq_mlp_infer_bf16_masked_async
    fm.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    fm.restype = ctypes.c_int
    fs = L.hyq_mlp_infer_bf16_segments_async
    fs.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                   ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]
    fs.restype = ctypes.c_int
    assert fm(m.handle, Xb.data_ptr(), n, Pref.data_ptr(), None) == 0
    assert fs(m.handle, Xb.data_ptr(), Id.data_ptr(), Sd.data_ptr(), nseg, srows, Pc.data_ptr(), None) == 0
    torch.cuda.synchronize()
    a, b = Pref.cpu().numpy(), Pc.cpu().numpy()
    listed = (np.arange(n) % srows) < counts[np.arange(n) // srows]
    assert listed.sum() == counts.sum()
    assert np.array_equal(b[idx[listed]], a[listed]), np.abs(b[idx[listed]] - a[listed]).max()
    untouched = np.ones(nplanes, bool)
    untouched[idx[listed]] = False
    assert (b[untouched] == 7.0).all()


def test_width128_network_in_the_production_step():
    """The warm-start configuration: the 15 -> 128 -> 128 -> 128 -> 2 network
    (hyqmom.jl_amd/models/warmstart_mlp_h128.f32, trained on the GPU by the library's own
    kernels on 10^7 planes of crossing-jets runs) applied in the 120x120x60 step (Ma=1.5,
    Kn=0.01) from the state of step 40 (the training's held-out state): fewer Laguerre iterations than the analytic start, the
    same fallback cells, and the step within 1e-6 of each field of the plain FP64 step on
    every cell whose FP64 step is not chaotic under one-ulp input changes."""
    import hyqmom_amd as H
    import parity_util as PU
    from hyqmom_amd import _lib
    from hyqmom_amd.warmstart import MODEL_H128, WarmStartMLP
    N = (120, 120, 60)
    p = dict(_params(N, 1.5, 0.01, "fp64"), tmax=1e9, nnmax=10 ** 9)
    s = H.Solver(p)
    s.set_state(H.initial_state(p))
    s.run(1e9, 40)
    M = np.asfortranarray(s.get_state())
    s.close()

    def one(X, prec, model=None):
        q = dict(p, precision=prec)
        if model is not None:
            q["warmstart_model"] = model
        s = H.Solver(q)
        s.set_state(X)
        s.step()
        fl = s.flags()
        out = (s.get_state().reshape(-1, 35, order="F"), fl)
        s.close()
        return out

    net = WarmStartMLP.load(MODEL_H128)
    assert net.hidden == 128
    f64, _ = one(M, "fp64")
    new, fl_n = one(M, "mixed_newton")
    lrn, fl_l = one(M, "mixed_learned", net)
    it_n, it_l = _lib.newton_iterations(fl_n).mean(), _lib.newton_iterations(fl_l).mean()
    fb_n, fb_l = int(((fl_n >> 12) & 1).sum()), int(((fl_l >> 12) & 1).sum())
    rng = np.random.default_rng(61)
    sens = np.zeros(len(f64))
    for _ in range(2):
        sens = np.maximum(sens, PU.cell_err(one(PU.ulp_perturb(M, rng), "fp64")[0], f64, PU.field_scale(f64)))
    keep = sens <= PU.CHAOTIC

    def fields(X):
        r = X[:, 0]
        return np.stack([r, X[:, 1] / r, X[:, 5] / r, X[:, 15] / r, 0.5 * (X[:, 2] + X[:, 9] + X[:, 19])], axis=1)

    fo, fg = fields(f64)[keep], fields(lrn)[keep]
    err = (np.abs(fg - fo) / np.abs(fo).max(axis=0)).max(axis=0)
    print("width 128: Laguerre iterations per cell %.2f (analytic start %.2f, %.0f%% fewer), fallback "
          "cells %d / %d, per-field error vs FP64 %s over %d cells (%d chaotic)"
          % (it_l, it_n, 100 * (1 - it_l / it_n), fb_l, fb_n, ["%.1e" % e for e in err], keep.sum(), (~keep).sum()))
    assert it_l < 0.9 * it_n
    assert fb_l == fb_n
    assert np.all(err <= 1e-6), err
