"""Device-array helpers for the GPU tests (torch allocates; the C ABI computes)."""
import numpy as np


def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("GPU test needs a HIP device")
    return torch


def to_dev(a):
    torch = torch_cuda()
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def empty_dev(shape, dtype="float64"):
    torch = torch_cuda()
    return torch.empty(shape, dtype=getattr(torch, dtype), device="cuda")


def to_host(t):
    return t.cpu().numpy()


def rel_err(a, b, scale):
    """|a-b| / scale, scale broadcast per field."""
    return np.abs(np.asarray(a) - np.asarray(b)) / scale
