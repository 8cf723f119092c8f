"""Device memory and stream plumbing (PyTorch-ROCm tensors as HBM buffers)."""

from __future__ import annotations

import numpy as np
import torch


def require_gpu() -> torch.device:
    """The HIP path is the only path: fail loudly without a GPU."""
    if not torch.cuda.is_available():
        raise RuntimeError("jamma_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr(stream: torch.cuda.Stream | None = None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    return int(t.data_ptr())


def to_dev(a, dtype=None, device=None) -> torch.Tensor:
    device = device or require_gpu()
    if isinstance(a, torch.Tensor):
        t = a.to(device=device, dtype=dtype or a.dtype)
    else:
        arr = np.ascontiguousarray(a)
        t = torch.from_numpy(arr).to(device=device)
        if dtype is not None:
            t = t.to(dtype)
    return t.contiguous()


def empty(shape, dtype, device=None) -> torch.Tensor:
    return torch.empty(shape, dtype=dtype, device=device or require_gpu())


def zeros(shape, dtype, device=None) -> torch.Tensor:
    return torch.zeros(shape, dtype=dtype, device=device or require_gpu())


def round_up(x: int, m: int) -> int:
    return ((x + m - 1) // m) * m
