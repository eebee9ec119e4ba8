"""The learned warm start of the wave-speed solve: a 15 -> H -> H -> H -> 2 GELU MLP (H = 32,
the shipped network, or 128) on bf16 MFMA that predicts the standardized extreme eigenvalues of
each plane's flux Jacobian (no counterpart in HyQMOM.jl).  Everything runs in libhyqmom_hip.so; this module
holds device buffers (torch tensors, plumbing only) and the training loop.

Precision "mixed_learned" uses the predictions as starting points of Laguerre's iteration,
which converges to the FP64 extreme roots (a start is used only when a Sturm count proves it
lies outside all roots) or the cell falls back to the full eigen solve, so the network
changes the iteration count, never the answer beyond the iteration's tolerance.
"""
import ctypes
import os

import numpy as np

from . import _lib

MODEL_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "models")
DEFAULT_MODEL = os.path.join(MODEL_DIR, "warmstart_mlp.f32")
# the warm-start configuration's network: three hidden layers of width 128
MODEL_H128 = os.path.join(MODEL_DIR, "warmstart_mlp_h128.f32")


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.HyqError("the warm-start network needs a HIP device")
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _sync():
    # the library works on its own non-blocking streams: inputs written by torch must be
    # complete before a call
    _torch().cuda.current_stream().synchronize()


HIDDEN_WIDTHS = (32, 128)


def width_of_blob(nblob):
    """The hidden width of a parameter blob of nblob floats."""
    L = _lib.lib()
    for H in HIDDEN_WIDTHS:
        if int(L.hyq_mlp_blob_size_hidden(H)) == nblob:
            return H
    raise ValueError("no supported hidden width has a %d-float blob" % nblob)


class WarmStartMLP:
    def __init__(self, blob=None, hidden=None):
        """A network from a parameter blob (its width follows from its size) or, with no blob,
        zeros of the given hidden width (32 by default)."""
        L = _lib.lib()
        if blob is not None:
            blob = np.ascontiguousarray(blob, dtype=np.float32)
            H = width_of_blob(blob.size)
            if hidden is not None and hidden != H:
                raise ValueError("blob is for hidden width %d, not %d" % (H, hidden))
        else:
            H = 32 if hidden is None else int(hidden)
            if H not in HIDDEN_WIDTHS:
                raise ValueError("hidden width must be one of %s" % (HIDDEN_WIDTHS,))
        self.hidden = H
        self.nblob = int(L.hyq_mlp_blob_size_hidden(H))
        self.nparam = int(L.hyq_mlp_param_count_hidden(H))
        h = ctypes.c_void_p()
        ptr = blob.ctypes.data_as(ctypes.c_void_p) if blob is not None else None
        rc = L.hyq_mlp_create_hidden(H, ptr, ctypes.byref(h))
        self._check(rc, "hyq_mlp_create_hidden")
        self.handle = h

    @staticmethod
    def _check(rc, what):
        if rc != 0:
            raise _lib.HyqError("%s failed: %s (%s)" % (what, _lib.ERRORS.get(rc, rc),
                                                        _lib.lib().hyq_mlp_last_error().decode()))

    @classmethod
    def load(cls, path=DEFAULT_MODEL):
        return cls(np.fromfile(path, dtype=np.float32))

    def save(self, path=DEFAULT_MODEL):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        self.blob().tofile(path)

    def blob(self):
        out = np.empty(self.nblob, np.float32)
        self._check(_lib.lib().hyq_mlp_get_blob(self.handle, out.ctypes.data_as(ctypes.c_void_p)),
                    "hyq_mlp_get_blob")
        return out

    def set_blob(self, blob):
        blob = np.ascontiguousarray(blob, dtype=np.float32)
        self._check(_lib.lib().hyq_mlp_set_blob(self.handle, blob.ctypes.data_as(ctypes.c_void_p)),
                    "hyq_mlp_set_blob")

    def infer(self, X):
        """X: device float32 [n][16] -> device float32 [n][2]."""
        torch = _torch()
        X = X.contiguous()
        _sync()
        P = torch.empty((X.shape[0], 2), dtype=torch.float32, device=X.device)
        self._check(_lib.lib().hyq_mlp_infer(self.handle, _ptr(X), X.shape[0], _ptr(P)), "hyq_mlp_infer")
        return P

    def set_shape(self, shape):
        """Fused inference launch shape: -1 picks the fastest by a timed trial, 0 (16 waves x 1 tile), 1 (8 x 2) or 2 (8 x 1) fixes one."""
        self._check(_lib.lib().hyq_mlp_set_shape(self.handle, int(shape)), "hyq_mlp_set_shape")

    def infer_layers(self, X):
        """The layer-by-layer path (training kernels) for the same prediction."""
        torch = _torch()
        X = X.contiguous()
        _sync()
        P = torch.empty((X.shape[0], 2), dtype=torch.float32, device=X.device)
        self._check(_lib.lib().hyq_mlp_infer_layers(self.handle, _ptr(X), X.shape[0], _ptr(P)),
                    "hyq_mlp_infer_layers")
        return P

    def grad(self, X, T):
        loss = ctypes.c_float()
        g = np.empty(self.nparam, np.float32)
        X, T = X.contiguous(), T.contiguous()
        _sync()
        self._check(_lib.lib().hyq_mlp_grad(self.handle, _ptr(X), _ptr(T),
                                            X.shape[0], ctypes.byref(loss),
                                            g.ctypes.data_as(ctypes.c_void_p)), "hyq_mlp_grad")
        return float(loss.value), g

    def train_step(self, X, T, lr):
        loss = ctypes.c_float()
        X, T = X.contiguous(), T.contiguous()
        _sync()
        self._check(_lib.lib().hyq_mlp_train_step(self.handle, _ptr(X), _ptr(T), X.shape[0],
                                                  ctypes.c_float(lr), ctypes.byref(loss)),
                    "hyq_mlp_train_step")
        return float(loss.value)

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().hyq_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plane_dataset(M_soa):
    """Features X [5n][16], labels L [5n][2] and validity [5n] of the five plane Jacobians of
    the n states M_soa (device float64 [35][n])."""
    torch = _torch()
    n = M_soa.shape[1]
    X = torch.empty((5 * n, 16), dtype=torch.float32, device=M_soa.device)
    Lb = torch.empty((5 * n, 2), dtype=torch.float32, device=M_soa.device)
    V = torch.empty((5 * n,), dtype=torch.uint8, device=M_soa.device)
    M_soa = M_soa.contiguous()
    _sync()
    _lib.check(_lib.lib().hyq_plane_dataset(_ptr(M_soa), n, _ptr(X), _ptr(Lb), _ptr(V)),
               "hyq_plane_dataset")
    return X, Lb, V


def fit_normalization(X, L):
    """Blob normalization block (xmu, xsd, ymu, ysd) from valid training samples."""
    Xh = X.double().cpu().numpy()
    Lh = L.double().cpu().numpy()
    xmu = np.zeros(16)
    xsd = np.ones(16)
    xmu[:15] = Xh[:, :15].mean(0)
    sd = Xh[:, :15].std(0)
    xsd[:15] = np.where(sd > 1e-6, sd, 1.0)
    ymu = Lh.mean(0)
    ysd = np.maximum(Lh.std(0), 1e-6)
    return np.concatenate([xmu, xsd, ymu, ysd]).astype(np.float32)


def hidden_width(nparam):
    """The hidden width H of a network with nparam parameters, 2 H^2 + 35 H + 16 of them
    (W1 16 x H, three biases of H, W2 and W3 H x H, W4 H x 16, b4 16)."""
    H = int(round((-35 + np.sqrt(35 * 35 + 8 * (nparam - 16))) / 4))
    if 2 * H * H + 35 * H + 16 != nparam:
        raise ValueError("no hidden width has %d parameters" % nparam)
    return H


def init_blob(nparam, norm, seed=0):
    """He-initialized weights (output columns >= 2 zero) with the given normalization."""
    rng = np.random.default_rng(seed)
    H = hidden_width(nparam)
    sizes = [(16, H), (H,), (H, H), (H,), (H, H), (H,), (H, 16), (16,)]
    parts = []
    for s in sizes:
        if len(s) == 2:
            w = rng.standard_normal(s) * np.sqrt(2.0 / s[0])
            if s[1] == 16:
                w[:, 2:] = 0.0
            parts.append(w.ravel())
        else:
            parts.append(np.zeros(s))
    blob = np.concatenate(parts + [norm]).astype(np.float32)
    assert blob.size == nparam + 36
    return blob


def train(X, L, steps=3000, batch=8192, lr=2e-3, lr_final=1e-4, seed=0, log=None, hidden=32):
    """Fit a fresh network of the given hidden width to (X, L) (device tensors of valid
    samples).  Adam with a cosine learning-rate decay from lr to lr_final.  Returns (model,
    list of (step, loss))."""
    torch = _torch()
    L_ = _lib.lib()
    norm = fit_normalization(X, L)
    m = WarmStartMLP(init_blob(int(L_.hyq_mlp_param_count_hidden(hidden)), norm, seed))
    g = torch.Generator(device=X.device)
    g.manual_seed(seed)
    n = X.shape[0]
    hist = []
    for it in range(steps):
        idx = torch.randint(0, n, (min(batch, n),), device=X.device, generator=g)
        xb = X[idx].contiguous()
        tb = L[idx].contiguous()
        lr_t = lr_final + 0.5 * (lr - lr_final) * (1.0 + np.cos(np.pi * it / max(1, steps - 1)))
        loss = m.train_step(xb, tb, lr_t)
        if it % 100 == 0 or it == steps - 1:
            hist.append((it, loss))
            if log:
                log("step %d loss %.4e lr %.2e" % (it, loss, lr_t))
    return m, hist
