"""CPU reference of the warm-start network (test infrastructure only).

The network has no counterpart in HyQMOM.jl: it only seeds the root-polishing iteration of
the wave speeds, whose starting point the FP64 step verifies (Sturm count) and otherwise
replaces by the analytic bound.  This module restates, in numpy, what the device code computes:

  Xn = (X[:, :15] - xmu) / xsd, padded with a zero 16th column
  H1 = g(Xn W1 + b1), H2 = g(H1 W2 + b2), H3 = g(H2 W3 + b3), Y = H3 W4 + b4
  prediction (lmin, lmax) = Y[:, :2] * ysd + ymu
  loss = 1/(2n) sum_b sum_{o<2} (Y[b, o] - (T[b, o] - ymu[o]) / ysd[o])^2

with every matrix-product operand rounded to bfloat16 (round to nearest even) and products
accumulated in high precision, as the MFMA v_mfma_f32_16x16x32_bf16 instructions do (FP32
accumulate).  g is the tanh form of GELU, z sigmoid(2 sqrt(2/pi) (z + 0.044715 z^3)).
With bf16=False everything stays in FP64, which the finite-difference check of the backward
pass uses.

Parameter blob (FP32), hidden width H (32 or 128; the blob's size gives it): W1 [16][H],
b1 [H], W2 [H][H], b2, W3 [H][H], b3, W4 [H][16], b4 [16], then xmu [16], xsd [16], ymu [2],
ysd [2].
"""
import math

import numpy as np

IN, HID, OUT = 16, 32, 16


def shapes(H=HID):
    return [("W1", (IN, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)),
            ("W3", (H, H)), ("b3", (H,)), ("W4", (H, OUT)), ("b4", (OUT,))]


def nparam(H=HID):
    return sum(int(np.prod(s)) for _, s in shapes(H))


def nblob(H=HID):
    return nparam(H) + 16 + 16 + 2 + 2


SHAPES = shapes()
NPARAM = nparam()
NBLOB = nblob()


def width(blob):
    """The hidden width of a parameter blob."""
    for H in (32, 64, 96, 128):
        if nblob(H) == np.asarray(blob).size:
            return H
    raise ValueError("no hidden width has a %d-float blob" % np.asarray(blob).size)


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(blob):
    blob = np.asarray(blob)
    out, o = {}, 0
    for name, shp in shapes(width(blob)):
        n = int(np.prod(shp))
        out[name] = blob[o:o + n].reshape(shp)
        o += n
    out["xmu"] = blob[o:o + 16]
    out["xsd"] = blob[o + 16:o + 32]
    out["ymu"] = blob[o + 32:o + 34]
    out["ysd"] = blob[o + 34:o + 36]
    return out


def random_blob(seed=0, scale=1.0, hidden=HID):
    """He-style random weights with unit normalization statistics."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shp in shapes(hidden):
        if name.startswith("W"):
            fan_in = shp[0]
            parts.append(rng.standard_normal(shp).ravel() * scale * math.sqrt(2.0 / fan_in))
        else:
            parts.append(0.1 * scale * rng.standard_normal(shp))
    parts.append(0.2 * rng.standard_normal(16))          # xmu
    parts.append(1.0 + 0.5 * rng.random(16))             # xsd
    parts.append(np.array([-2.0, 2.0]))                  # ymu
    parts.append(np.array([0.5, 0.5]))                   # ysd
    blob = np.concatenate([np.ravel(p) for p in parts]).astype(np.float32)
    p = unpack(blob)
    p["W4"][:, 2:] = 0.0
    p["b4"][2:] = 0.0
    return blob


def _mm(A, B, use_bf16):
    if use_bf16:
        return bf16(A).astype(np.float64) @ bf16(B).astype(np.float64)
    return np.asarray(A, np.float64) @ np.asarray(B, np.float64)


def _sig(u):
    with np.errstate(over="ignore"):
        return 1.0 / (1.0 + np.exp(-u))


def _gelu(z):
    return z * _sig(1.5957691216057308 * (z + 0.044715 * z ** 3))


def _dgelu(z):
    s = _sig(1.5957691216057308 * (z + 0.044715 * z ** 3))
    return s + z * s * (1.0 - s) * 1.5957691216057308 * (1.0 + 0.134145 * z * z)


def _f(x, use_bf16):
    # activations are stored in FP32 on the device between layers
    return np.asarray(x, np.float32) if use_bf16 else np.asarray(x, np.float64)


def normalize(blob, X):
    p = unpack(blob)
    X = np.asarray(X, np.float64)
    Xn = np.zeros((X.shape[0], IN))
    Xn[:, :15] = (X[:, :15] - p["xmu"][:15]) / p["xsd"][:15]
    return Xn


def forward(blob, X, use_bf16=True):
    """Returns a dict of Xn, Z1, H1, Z2, H2, Z3, H3, Y (n rows each)."""
    p = unpack(blob)
    Xn = _f(normalize(blob, X), use_bf16)
    act = {"Xn": Xn}
    H = Xn
    for k in (1, 2, 3):
        Z = _mm(H, p["W%d" % k], use_bf16) + np.asarray(p["b%d" % k], np.float64)
        H = _f(_gelu(Z), use_bf16)
        act["Z%d" % k], act["H%d" % k] = _f(Z, use_bf16), H
    act["Y"] = _f(_mm(H, p["W4"], use_bf16) + np.asarray(p["b4"], np.float64), use_bf16)
    return act


def predict(blob, X, use_bf16=True):
    p = unpack(blob)
    Y = np.asarray(forward(blob, X, use_bf16)["Y"], np.float64)
    return Y[:, :2] * p["ysd"] + p["ymu"]


def loss_and_grad(blob, X, T, use_bf16=True):
    """(loss, gradient [NPARAM]) of the mean-squared loss above."""
    p = unpack(blob)
    act = forward(blob, X, use_bf16)
    n = X.shape[0]
    Y = np.asarray(act["Y"], np.float64)
    Tn = (np.asarray(T, np.float64) - p["ymu"]) / p["ysd"]
    d = Y[:, :2] - Tn
    loss = 0.5 * np.sum(d * d) / n
    dY = np.zeros((n, OUT))
    dY[:, :2] = d / n
    dY = _f(dY, use_bf16)
    g = {}
    g["W4"] = _mm(act["H3"].T, dY, use_bf16)
    g["b4"] = np.sum(np.asarray(dY, np.float64), axis=0)
    dA = _mm(dY, p["W4"].T, use_bf16) * _dgelu(np.asarray(act["Z3"], np.float64))
    dA = _f(dA, use_bf16)
    for k, hin in ((3, "H2"), (2, "H1"), (1, "Xn")):
        g["W%d" % k] = _mm(act[hin].T, dA, use_bf16)
        g["b%d" % k] = np.sum(np.asarray(dA, np.float64), axis=0)
        if k > 1:
            dA = _mm(dA, p["W%d" % k].T, use_bf16) * _dgelu(np.asarray(act["Z%d" % (k - 1)], np.float64))
            dA = _f(dA, use_bf16)
    grad = np.concatenate([np.ravel(g[name]) for name, _ in shapes(width(blob))])
    return loss, grad
