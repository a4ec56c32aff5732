"""Forced Taylor-Green MMS harness of the reference's
packages/lattice-boltzmann-d3q19/tests/test_mms_convergence.py:75-131, shared
by the oracle and GPU tests; it lives in the package as a verification tool."""

from lattice_boltzmann_d3q19_amd.verification import (  # noqa: F401
    KNOWN_ANSWERS,
    mms_error,
    mms_error_lattice,
    observed_order,
    setup,
    source,
)
