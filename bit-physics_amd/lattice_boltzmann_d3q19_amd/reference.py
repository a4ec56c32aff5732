"""Drop-in for ``lattice_boltzmann_d3q19.reference`` computed on the GPU.

Same names, argument meaning and return conventions as the reference
surface (packages/lattice-boltzmann-d3q19/lattice_boltzmann_d3q19/reference/__init__.py:43-71):
functions take ``(19, Nx, Ny, Nz)`` population arrays, never mutate their
inputs and return new arrays.  NumPy in -> NumPy float64 out; a torch
device tensor in -> a device tensor out (no host round trip).

Every call runs in libd3q19 (HIP, gfx950).  ``precision="fp64"`` (the
default, matching the reference's float64) uses kernels that follow the
reference's operation order with FMA contraction off; ``"fp32"`` runs the
same kernels in single precision.
"""

from __future__ import annotations

import numpy as np

from . import _device as dev
from ._lib import D3Q19_F32, D3Q19_F64, call

# ---- constants (reference/constants.py:39-100) ---------------------------
VELOCITIES = (
    (0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
    (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1),
    (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1),
)
WEIGHTS = (1.0 / 3.0, *([1.0 / 18.0] * 6), *([1.0 / 36.0] * 12))
CS2 = 1.0 / 3.0
C = np.asarray(VELOCITIES, dtype=np.int64)
W = np.asarray(WEIGHTS, dtype=np.float64)
OPP = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17], dtype=np.int64)

CANONICAL_DESCRIPTOR_POISEUILLE = "poiseuille-64x32-seed42-step1000"
CANONICAL_DESCRIPTOR_COUETTE = "couette-32x16-seed42-step500"
CANONICAL_SEED = 42
CANONICAL_NZ = 3
CANONICAL_POISEUILLE_NX = 64
CANONICAL_POISEUILLE_NY = 32
CANONICAL_POISEUILLE_STEPS = 1000
CANONICAL_COUETTE_NX = 32
CANONICAL_COUETTE_NY = 16
CANONICAL_COUETTE_STEPS = 500

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


# ---- equilibrium.py ------------------------------------------------------
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
    """equilibrium.py:34-53: the 19 f_i^eq at one (rho, u)."""
    r = np.full((1, 1, 1), float(rho))
    uu = np.asarray([float(u[0]), float(u[1]), float(u[2])], dtype=np.float64).reshape(3, 1, 1, 1)
    return [float(v) for v in feq_field(r, uu, precision=precision).reshape(19)]


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


# ---- bgk.py --------------------------------------------------------------
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
    import ctypes

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


__all__ = [
    "C", "CANONICAL_COUETTE_NX", "CANONICAL_COUETTE_NY", "CANONICAL_COUETTE_STEPS",
    "CANONICAL_DESCRIPTOR_COUETTE", "CANONICAL_DESCRIPTOR_POISEUILLE", "CANONICAL_NZ",
    "CANONICAL_POISEUILLE_NX", "CANONICAL_POISEUILLE_NY", "CANONICAL_POISEUILLE_STEPS",
    "CANONICAL_SEED", "CS2", "VELOCITIES", "W", "WEIGHTS", "apply_bounce_back_y_walls",
    "bgk_step", "density_field", "density_moment", "feq", "feq_field",
    "macroscopic_velocity", "momentum_field", "momentum_moment", "stream",
]
