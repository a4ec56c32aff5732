"""Drop-in for ``lattice_boltzmann_d3q19.reference`` computed on the GPU.

Same names, submodules (``bgk``, ``equilibrium``, ``constants``), argument
meaning and return conventions as the reference surface
(packages/lattice-boltzmann-d3q19/lattice_boltzmann_d3q19/reference/__init__.py:10-71):
functions take ``(19, Nx, Ny, Nz)`` population arrays, never mutate their
inputs and return new arrays.  NumPy in -> NumPy float64 out; a torch
device tensor in -> a device tensor out (no host round trip).

Every call runs in libd3q19 (HIP, gfx950).  ``precision="fp64"`` (the
default, matching the reference's float64) uses kernels that follow the
reference's operation order with FMA contraction off; ``"fp32"`` runs the
same kernels in single precision.
"""

from __future__ import annotations

from . import bgk, constants, equilibrium
from .bgk import apply_bounce_back_y_walls, bgk_step, macroscopic_velocity, stream
from .constants import (
    CANONICAL_COUETTE_NX,
    CANONICAL_COUETTE_NY,
    CANONICAL_COUETTE_STEPS,
    CANONICAL_DESCRIPTOR_COUETTE,
    CANONICAL_DESCRIPTOR_POISEUILLE,
    CANONICAL_NZ,
    CANONICAL_POISEUILLE_NX,
    CANONICAL_POISEUILLE_NY,
    CANONICAL_POISEUILLE_STEPS,
    CANONICAL_SEED,
    CS2,
    OPP,
    VELOCITIES,
    WEIGHTS,
    C,
    W,
)
from .equilibrium import (
    density_field,
    density_moment,
    feq,
    feq_field,
    momentum_field,
    momentum_moment,
)

__all__ = [
    "C", "CANONICAL_COUETTE_NX", "CANONICAL_COUETTE_NY", "CANONICAL_COUETTE_STEPS",
    "CANONICAL_DESCRIPTOR_COUETTE", "CANONICAL_DESCRIPTOR_POISEUILLE", "CANONICAL_NZ",
    "CANONICAL_POISEUILLE_NX", "CANONICAL_POISEUILLE_NY", "CANONICAL_POISEUILLE_STEPS",
    "CANONICAL_SEED", "CS2", "VELOCITIES", "W", "WEIGHTS", "apply_bounce_back_y_walls",
    "bgk", "bgk_step", "density_field", "density_moment", "equilibrium", "feq", "feq_field",
    "macroscopic_velocity", "momentum_field", "momentum_moment", "stream",
]
