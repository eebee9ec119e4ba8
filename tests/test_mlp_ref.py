"""The CPU reference of the warm-start network: bf16 rounding and the backward pass, pinned
by known values and central finite differences of the FP64 forward pass."""
import numpy as np
import pytest

import mlp_ref as R


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -7 + 2.0 ** -9,
                  -2.5, 3.0e-39, 65504.0], dtype=np.float32)
    # 1+2^-8 is a tie -> even (1.0); 1+3*2^-8 is a tie -> even (1+2^-6);
    # 1+2^-7+2^-9 rounds to 1+2^-7
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7, -2.5, 0.0, 65536.0], dtype=np.float32)
    got = R.bf16(x)
    assert got[:5].tolist() == want[:5].tolist()
    assert abs(got[5]) < 3.1e-39  # subnormal keeps its top bits
    assert got[6] == want[6]


def test_blob_layout_sizes():
    assert R.HID == 32
    assert R.NPARAM == 16 * 32 + 32 + 2 * (32 * 32 + 32) + 32 * 16 + 16
    assert R.NBLOB == R.NPARAM + 36
    p = R.unpack(R.random_blob(1))
    assert np.all(p["W4"][:, 2:] == 0) and np.all(p["b4"][2:] == 0)


def test_gradient_matches_finite_differences():
    blob = R.random_blob(3).astype(np.float64)
    rng = np.random.default_rng(4)
    X = rng.standard_normal((7, 16))
    T = rng.standard_normal((7, 2))
    _, g = R.loss_and_grad(blob, X, T, use_bf16=False)
    # parameters of every tensor, including the first and last of each
    idx, o = [], 0
    for _, shp in R.SHAPES:
        n = int(np.prod(shp))
        idx += [o, o + n - 1] + list(rng.integers(o, o + n, 6))
        o += n
    # columns >= 2 of W4/b4 do not reach the loss
    p = R.unpack(np.arange(R.NBLOB, dtype=np.float64))
    dead = set(p["W4"][:, 2:].astype(int).ravel()) | set(p["b4"][2:].astype(int))
    for i in idx:
        h = 1e-6 * max(1.0, abs(blob[i]))
        bp, bm = blob.copy(), blob.copy()
        bp[i] += h
        bm[i] -= h
        fd = (R.loss_and_grad(bp, X, T, False)[0] - R.loss_and_grad(bm, X, T, False)[0]) / (2 * h)
        if i in dead:
            assert g[i] == 0.0 and abs(fd) < 1e-12
        else:
            assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(fd)), (i, fd, g[i])


def test_bf16_forward_close_to_fp64():
    blob = R.random_blob(5)
    X = np.random.default_rng(6).standard_normal((64, 16))
    a = R.predict(blob, X, True)
    b = R.predict(blob, X, False)
    # bf16 operands: ~2^-9 relative per operand, accumulated over 3 hidden layers
    assert np.max(np.abs(a - b)) < 0.05 * max(1.0, np.max(np.abs(b)))
    assert not np.array_equal(a, b)


@pytest.mark.parametrize("n", [1, 5])
def test_loss_zero_at_exact_labels(n):
    blob = R.random_blob(7)
    X = np.random.default_rng(8).standard_normal((n, 16))
    T = R.predict(blob, X, False)
    loss, g = R.loss_and_grad(blob, X, T, False)
    assert loss < 1e-20 and np.max(np.abs(g)) < 1e-9
