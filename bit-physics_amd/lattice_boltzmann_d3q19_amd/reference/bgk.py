"""BGK step, streaming, velocity and y-wall bounce-back on the GPU (reference/bgk.py)."""

from __future__ import annotations

import ctypes

from .. import _device as dev
from .._lib import call
from ._common import _dt, _field3, _out, _pop_shape

def bgk_step(f, tau: float, force_lattice=None, *, precision: str = "fp64"):
    """bgk.py:45-112: BGK collision (+ Guo force) then periodic streaming."""
    code, tdt = _dt(precision)
    shape = _pop_shape(f)
    _field3(force_lattice, shape[1:], "force_lattice")
    ff = dev.to_device(f, tdt)
    F = None if force_lattice is None else dev.to_device(force_lattice, tdt)
    out = dev.empty(shape, tdt)
    nx, ny, nz = shape[1:]
    call("d3q19_bgk_step", dev.ptr(ff), dev.ptr(out), dev.ptr(F), nx, ny, nz, float(tau), code,
         dev.stream_ptr())
    return _out(out, f)


def stream(f_post, *, precision: str = "fp64"):
    """bgk.py:115-126: out[i, x] = f_post[i, x - c_i] (periodic)."""
    code, tdt = _dt(precision)
    shape = _pop_shape(f_post)
    ff = dev.to_device(f_post, tdt)
    out = dev.empty(shape, tdt)
    nx, ny, nz = shape[1:]
    call("d3q19_stream", dev.ptr(ff), dev.ptr(out), nx, ny, nz, code, dev.stream_ptr())
    return _out(out, f_post)


def macroscopic_velocity(f, force_lattice=None, *, precision: str = "fp64"):
    """bgk.py:129-142: u = (rho u + F/2) / max(rho, 1e-30)."""
    code, tdt = _dt(precision)
    shape = _pop_shape(f)
    _field3(force_lattice, shape[1:], "force_lattice")
    ff = dev.to_device(f, tdt)
    F = None if force_lattice is None else dev.to_device(force_lattice, tdt)
    out = dev.empty((3,) + shape[1:], tdt)
    call("d3q19_macroscopic_velocity", dev.ptr(ff), dev.ptr(F), dev.ptr(out), out[0].numel(), code,
         dev.stream_ptr())
    return _out(out, f)


def apply_bounce_back_y_walls(f, wall_velocity_top=(0.0, 0.0, 0.0),
                              wall_velocity_bottom=(0.0, 0.0, 0.0), *, precision: str = "fp64"):
    """bgk.py:145-216: y-wall bounce-back with moving-wall momentum term."""
    code, tdt = _dt(precision)
    shape = _pop_shape(f)
    ff = dev.to_device(f, tdt)
    out = dev.empty(shape, tdt)
    top = (ctypes.c_double * 3)(*[float(v) for v in wall_velocity_top])
    bot = (ctypes.c_double * 3)(*[float(v) for v in wall_velocity_bottom])
    nx, ny, nz = shape[1:]
    call("d3q19_bounce_back_y_walls", dev.ptr(ff), dev.ptr(out), top, bot, nx, ny, nz, code,
         dev.stream_ptr())
    return _out(out, f)


__all__ = ["apply_bounce_back_y_walls", "bgk_step", "macroscopic_velocity", "stream"]
