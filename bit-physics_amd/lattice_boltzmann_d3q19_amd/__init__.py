"""MI355X-native D3Q19 lattice-Boltzmann engine (HIP, gfx950).

Drop-in for the step / read-back path of the reference package
``lattice_boltzmann_d3q19`` (packages/lattice-boltzmann-d3q19):

* :mod:`.reference` -- the reference's stateless functional API
  (``bgk_step``, ``stream``, ``apply_bounce_back_y_walls``, ``feq_field`` ...)
  computed on the GPU;
* :class:`.Lattice` -- the stateful step / read-back path (fused kernel);
* :mod:`.autograd` -- ``torch.autograd.Function`` over the adjoint kernels;
* :mod:`.initial` -- Taylor-Green and random-phase initial fields.

All computation runs in ``libd3q19.so`` through its C ABI (include/d3q19.h).
"""

from . import reference
from ._lib import LIB_PATH, D3Q19Error, load
from .lattice import Lattice

__all__ = ["LIB_PATH", "D3Q19Error", "Lattice", "load", "reference"]
__version__ = "0.1.0"
