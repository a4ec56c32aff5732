"""Initial fields of the studies (lattice_boltzmann_d3q19_amd.initial), on the CPU.

random_solenoidal must be divergence-free in the spectral sense for odd and
even extents alike (even extents have Nyquist modes, whose wavenumber sign is
ambiguous: they carry no velocity), hit the requested rms exactly, and repeat
for a seed.  The reference for "divergence-free" is the field's own discrete
Fourier transform: max |k . u_k| over max |k| |u_k|.
"""

from __future__ import annotations

import numpy as np
import pytest

from lattice_boltzmann_d3q19_amd.initial import random_solenoidal, spectral_divergence, taylor_green


@pytest.mark.parametrize("shape", [(16, 16, 16), (16, 12, 10), (15, 15, 15), (9, 16, 7), (8, 8, 8)])
def test_random_solenoidal_is_divergence_free(shape):
    u = random_solenoidal(shape, 0.03, seed=42)
    assert u.shape == (3,) + shape
    assert spectral_divergence(u) < 1e-13
    assert abs(np.sqrt(np.mean(np.sum(u * u, axis=0)) / 3.0) - 0.03) < 1e-15
    assert np.array_equal(u, random_solenoidal(shape, 0.03, seed=42))
    assert not np.array_equal(u, random_solenoidal(shape, 0.03, seed=43))


def test_spectral_divergence_sees_a_compressive_field():
    # u = grad(phi) for a single Fourier mode: divergence equals |k| |u_k|
    n = 12
    x = 2 * np.pi * np.arange(n) / n
    X, _, _ = np.meshgrid(x, x, x, indexing="ij")
    u = np.stack([np.sin(X), np.zeros_like(X), np.zeros_like(X)])
    assert spectral_divergence(u) > 0.99


def test_taylor_green_is_divergence_free():
    _, u = taylor_green((12, 12, 12), 0.04)
    assert spectral_divergence(u) < 1e-13
