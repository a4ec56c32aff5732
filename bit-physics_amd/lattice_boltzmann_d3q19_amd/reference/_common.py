"""Argument checks and host/device conversion shared by the drop-in functions."""

from __future__ import annotations

from .. import _device as dev
from .._lib import D3Q19_F32, D3Q19_F64, call

_PRECISION = {"fp64": D3Q19_F64, "fp32": D3Q19_F32}


def _dt(precision: str):
    try:
        code = _PRECISION[precision]
    except KeyError:
        raise ValueError(f"precision must be 'fp64' or 'fp32', got {precision!r}") from None
    t = dev.torch()
    return code, (t.float64 if code == D3Q19_F64 else t.float32)


def _out(result, like):
    """Return a device tensor for tensor input, a float64 NumPy array otherwise."""
    if dev.is_tensor(like):
        return result
    return result.to(dtype=dev.torch().float64).cpu().numpy()


def _pop_shape(f):
    shape = tuple(f.shape)
    if len(shape) != 4 or shape[0] != 19:
        raise ValueError(f"populations must have shape (19, Nx, Ny, Nz), got {shape}")
    return shape


def _field3(a, grid, name):
    if a is None:
        return None
    if tuple(a.shape) != (3,) + tuple(grid):
        raise ValueError(f"{name} must have shape (3, Nx, Ny, Nz) = {(3,) + tuple(grid)}, got {tuple(a.shape)}")
    return a


