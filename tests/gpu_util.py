"""Helpers for GPU tests: torch provides device memory and the stream only."""
import ctypes as C

import numpy as np
import pytest

try:
    import torch
except Exception:  # pragma: no cover - torch is part of the image
    torch = None


def require_gpu():
    if torch is None or not torch.cuda.is_available():
        pytest.skip("no GPU on this host")


def dev(a, dtype=None):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype or torch.float64, device="cuda")


def empty(n, dtype=None):
    return torch.empty(n, dtype=dtype or torch.float64, device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def stream():
    return torch.cuda.current_stream().cuda_stream


def point_table(P):
    """Device array of per-dimension device pointers for a (dim, n) tensor."""
    return torch.tensor([P[d].data_ptr() for d in range(P.shape[0])], dtype=torch.int64, device="cuda")


def workspace(nbytes):
    return torch.empty(max(int(nbytes), 8) // 8 + 1, dtype=torch.float64, device="cuda")
