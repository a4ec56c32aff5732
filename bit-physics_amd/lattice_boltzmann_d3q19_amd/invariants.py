"""Property-based equilibrium invariants, evaluated on the GPU drop-ins.

Same callables as the reference's gate-12 module
(packages/lattice-boltzmann-d3q19/lattice_boltzmann_d3q19/invariants.py:32-110):
each is a Hypothesis-decorated function of no arguments that samples
(rho, u) in the weakly-compressible band (rho in [0.5, 1.5], each |u_a|
below Ma * c_s with Ma = 0.1, samples with |u| / c_s >= 0.1 skipped) and
asserts the identity

* :func:`equilibrium_density_moment`: sum_i f_i^eq(rho, u) = rho,
* :func:`equilibrium_momentum_moment`: sum_i c_i f_i^eq(rho, u) = rho u,

to 1e-14 relative (floored at 1), with ``feq``, ``density_moment`` and
``momentum_moment`` computed by libd3q19's FP64 kernels.  Importing this
module needs ``hypothesis``; the rest of the package does not.
"""

from __future__ import annotations

from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

from .reference import CS2, density_moment, feq, momentum_moment

MA_BOUND = 0.1
U_BOUND = MA_BOUND * CS2 ** 0.5
DENSITY_TOL = 1e-14
MOMENTUM_TOL = 1e-14

_rho = st.floats(min_value=0.5, max_value=1.5, allow_nan=False, allow_infinity=False)
_u = st.floats(min_value=-U_BOUND, max_value=U_BOUND, allow_nan=False, allow_infinity=False)
_settings = settings(max_examples=50, deadline=None, suppress_health_check=[HealthCheck.function_scoped_fixture])


def _in_band(ux: float, uy: float, uz: float) -> bool:
    return (ux * ux + uy * uy + uz * uz) ** 0.5 / CS2 ** 0.5 < MA_BOUND


@given(rho=_rho, ux=_u, uy=_u, uz=_u)
@_settings
def equilibrium_density_moment(rho: float, ux: float, uy: float, uz: float) -> None:
    """sum_i f_i^eq(rho, u) = rho within 1e-14 * max(|rho|, 1)."""
    if not _in_band(ux, uy, uz):
        return
    got = density_moment(feq(rho, (ux, uy, uz)))
    assert abs(got - rho) <= DENSITY_TOL * max(abs(rho), 1.0), (
        f"equilibrium_density_moment violated: rho_in={rho:.6e} rho_out={got:.6e}"
    )


@given(rho=_rho, ux=_u, uy=_u, uz=_u)
@_settings
def equilibrium_momentum_moment(rho: float, ux: float, uy: float, uz: float) -> None:
    """sum_i c_i f_i^eq(rho, u) = rho u_a per component within 1e-14 * max(|rho u_a|, 1)."""
    if not _in_band(ux, uy, uz):
        return
    got = momentum_moment(feq(rho, (ux, uy, uz)))
    for a, (g, ua) in enumerate(zip(got, (ux, uy, uz))):
        want = rho * ua
        assert abs(g - want) <= MOMENTUM_TOL * max(abs(want), 1.0), (
            f"equilibrium_momentum_moment[{'xyz'[a]}] violated: expected={want:.6e} got={g:.6e}"
        )


__all__ = ["equilibrium_density_moment", "equilibrium_momentum_moment"]
