"""Equilibrium and moments on the GPU (reference/equilibrium.py)."""

from __future__ import annotations

import numpy as np

from .. import _device as dev
from .._lib import call
from ._common import _dt, _field3, _out, _pop_shape

def feq_field(rho, u, *, precision: str = "fp64"):
    """equilibrium.py:71-98: f_eq (19, Nx, Ny, Nz) from rho (Nx,Ny,Nz), u (3,Nx,Ny,Nz)."""
    code, tdt = _dt(precision)
    grid = tuple(rho.shape)
    _field3(u, grid, "u")
    r = dev.to_device(rho, tdt)
    uu = dev.to_device(u, tdt)
    out = dev.empty((19,) + grid, tdt)
    call("d3q19_feq_field", dev.ptr(r), dev.ptr(uu), dev.ptr(out), r.numel(), code, dev.stream_ptr())
    return _out(out, rho)


def feq(rho: float, u, *, precision: str = "fp64") -> list[float]:
    """equilibrium.py:34-53: the 19 f_i^eq at one (rho, u), in the point
    function's operation order (divisions by c_s^2, 2 c_s^4, 2 c_s^2)."""
    code, tdt = _dt(precision)
    r = dev.to_device(np.full(1, float(rho)), tdt)
    uu = dev.to_device(np.asarray([float(u[0]), float(u[1]), float(u[2])], dtype=np.float64).reshape(3, 1), tdt)
    out = dev.empty((19, 1), tdt)
    call("d3q19_feq", dev.ptr(r), dev.ptr(uu), dev.ptr(out), 1, code, dev.stream_ptr())
    return [float(v) for v in out.to(dtype=dev.torch().float64).cpu().numpy().reshape(19)]


def density_field(f, *, precision: str = "fp64"):
    """equilibrium.py:101-103."""
    code, tdt = _dt(precision)
    shape = _pop_shape(f)
    ff = dev.to_device(f, tdt)
    out = dev.empty(shape[1:], tdt)
    call("d3q19_density_field", dev.ptr(ff), dev.ptr(out), out.numel(), code, dev.stream_ptr())
    return _out(out, f)


def momentum_field(f, *, precision: str = "fp64"):
    """equilibrium.py:106-113."""
    code, tdt = _dt(precision)
    shape = _pop_shape(f)
    ff = dev.to_device(f, tdt)
    out = dev.empty((3,) + shape[1:], tdt)
    call("d3q19_momentum_field", dev.ptr(ff), dev.ptr(out), out[0].numel(), code, dev.stream_ptr())
    return _out(out, f)


def density_moment(f) -> float:
    """equilibrium.py:56-58 for one cell's 19 populations."""
    arr = np.asarray([float(v) for v in f], dtype=np.float64).reshape(19, 1, 1, 1)
    return float(density_field(arr).reshape(()))


def momentum_moment(f) -> list[float]:
    """equilibrium.py:61-68 for one cell's 19 populations."""
    arr = np.asarray([float(v) for v in f], dtype=np.float64).reshape(19, 1, 1, 1)
    return [float(v) for v in momentum_field(arr).reshape(3)]


__all__ = ["density_field", "density_moment", "feq", "feq_field", "momentum_field", "momentum_moment"]
