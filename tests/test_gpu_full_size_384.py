"""Decaying turbulence at its full size, 384^3 (tau = 0.51, random-phase
solenoidal u with rms 0.03, seed 42): rows of 384 cells run as three row
segments of 128, a launch geometry the 256^3 workload does not use.

* Two steps against the FP64 oracle (run_reference): FP64 storage field-scaled
  below 1e-12 in f - w; FP32 storage below 1e-6 relative in populations.
* Translation by (101, 7, 250) cells commutes with 10 steps, bit for bit, in
  FP32, FP16 and BF16 storage.
* The whole 4000-step run against FP64 storage: kinetic energy every 500
  steps within 1e-5 (FP32) and the 16-bit target 1e-3 (FP16, stochastically
  rounded BF16).
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import print_case

pytestmark = pytest.mark.gpu

SHAPE, TAU = (384, 384, 384), 0.51
SHIFT = (101, 7, 250)


@pytest.fixture(scope="module")
def mods():
    import torch

    import lattice_boltzmann_d3q19_amd as amd
    from lattice_boltzmann_d3q19_amd.initial import random_solenoidal

    return torch, amd, random_solenoidal(SHAPE, 0.03, seed=42)


def stepped(torch, amd, storage, u, steps):
    lat = amd.Lattice(SHAPE, TAU, storage=storage).set_equilibrium(None, u)
    lat.step(steps)
    f = torch.empty((19,) + SHAPE, dtype=torch.float64, device="cuda")
    lat.populations_into(f)
    lat.close()
    return f


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_two_steps_match_oracle(mods, storage):
    torch, amd, u0 = mods
    print_case("full-size-384-oracle", shape=SHAPE, storage=storage, steps=2)
    f = stepped(torch, amd, storage, u0, 2).cpu().numpy()
    ref = oracle.run_reference(oracle.feq_field(np.ones(SHAPE), u0), TAU, 2)
    _, w, _, _ = oracle.constants()
    wb = w[:, None, None, None]
    dev = np.abs((f - wb) - (ref - wb)).max() / np.abs(ref - wb).max()
    rel = np.abs(f - ref).max() / np.abs(ref).max()
    print(f"384^3 {storage}: relative in populations {rel:.3e}, field-scaled f - w {dev:.3e}", flush=True)
    del f, ref
    if storage == "f64":
        assert dev < 1e-12
    else:
        assert rel < 1e-6


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_translation_equivariance_bitwise(mods, storage):
    torch, amd, u0 = mods
    print_case("full-size-384-translation", shape=SHAPE, storage=storage, shift=SHIFT)
    u = torch.tensor(u0, device="cuda")
    fa = stepped(torch, amd, storage, u, 10)
    fb = stepped(torch, amd, storage, torch.roll(u, SHIFT, (1, 2, 3)), 10)
    assert torch.equal(torch.roll(fa, SHIFT, (1, 2, 3)), fb)
    del fa, fb, u
    torch.cuda.empty_cache()


def ke_series(amd, storage, u0, steps=4000, every=500):
    lat = amd.Lattice(SHAPE, TAU, storage=storage).set_equilibrium(None, u0)
    ke = [lat.diagnostics()["kinetic_energy"]]
    for _ in range(steps // every):
        ke.append(lat.step(every).diagnostics()["kinetic_energy"])
    lat.close()
    return np.array(ke)


@pytest.fixture(scope="module")
def ke_fp64(mods):
    _, amd, u0 = mods
    return ke_series(amd, "f64", u0)


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f16", 1e-3), ("bf16sr", 1e-3)])
def test_full_run_kinetic_energy_against_fp64(mods, ke_fp64, storage, tol):
    """The whole 4000-step run: kinetic energy every 500 steps against FP64
    storage.  The FP16 bound is the path's 16-bit target (1e-3 relative);
    measured FP32 5.8e-6, FP16 1.4e-4, BF16 with stochastic rounding 1.3e-4."""
    torch, amd, u0 = mods
    print_case("full-size-384-ke", shape=SHAPE, storage=storage, steps=4000)
    ke = ke_series(amd, storage, u0)
    err = float(np.max(np.abs(ke - ke_fp64) / ke_fp64))
    print(f"384^3 {storage}: max relative KE error over 4000 steps {err:.3e}", flush=True)
    assert err < tol
    assert ke_fp64[-1] < 0.2 * ke_fp64[0]  # the turbulence decays
