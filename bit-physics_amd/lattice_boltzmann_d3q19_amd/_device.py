"""Device-memory plumbing: NumPy <-> torch device tensors <-> raw pointers.

torch supplies device allocations and the current HIP stream; every
computation runs in libd3q19.
"""

from __future__ import annotations

import ctypes

import numpy as np

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as _t

        if not _t.cuda.is_available():
            raise RuntimeError("libd3q19 needs a ROCm GPU: torch.cuda.is_available() is False")
        _torch = _t
    return _torch


def is_tensor(a) -> bool:
    t = _torch
    if t is None:
        try:
            import torch as t  # noqa: F811
        except ImportError:
            return False
    return isinstance(a, t.Tensor)


def stream_ptr(device: int = 0) -> ctypes.c_void_p:
    t = torch()
    return ctypes.c_void_p(t.cuda.current_stream(device).cuda_stream)


def ptr(a) -> ctypes.c_void_p:
    """Raw device pointer of a torch tensor (None passes through as NULL)."""
    if a is None:
        return ctypes.c_void_p(None)
    return ctypes.c_void_p(a.data_ptr())


def to_device(a, dtype, device: int = 0):
    """NumPy array or tensor -> contiguous device tensor of `dtype`."""
    t = torch()
    if is_tensor(a):
        return a.to(device=f"cuda:{device}", dtype=dtype).contiguous()
    arr = np.ascontiguousarray(a)
    return t.from_numpy(arr).to(device=f"cuda:{device}", dtype=dtype).contiguous()


def empty(shape, dtype, device: int = 0):
    return torch().empty(tuple(shape), dtype=dtype, device=f"cuda:{device}")
