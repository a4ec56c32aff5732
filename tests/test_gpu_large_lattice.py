"""Lattices sized for one GPU's HBM, and the one-pass equilibrium load.

* set_equilibrium computes feq and stores the deviations in one kernel; the
  state equals, byte for byte, the state set_populations gives from
  feq_field(rho, u) (the reference's equilibrium.py:71-98), in every storage
  and for FP64 and FP32 inputs.
* A 1400 x 16 x 65536 lattice (1.47e9 cells, 3 * ncell > 2^32: the second and
  third components of (3, ncell) fields sit past 32-bit offsets) with FP16
  storage, a per-cell force field and an initial velocity that depend on y
  only.  The flow stays independent of x and z, so every (x, z) column must
  equal, bit for bit, the same run on a 3 x 16 x 64 lattice: a wrapped offset
  anywhere in the force read, the equilibrium load or the read-back would put
  values in the wrong cells.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from conftest import print_case

pytestmark = pytest.mark.gpu

STORAGES = ["f64", "f32", "f16", "bf16", "bf16sr"]


@pytest.fixture(scope="module")
def mods():
    import torch

    import lattice_boltzmann_d3q19_amd as amd

    return torch, amd


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("shape", [(6, 5, 4), (3, 4, 70)], ids=lambda s: "x".join(map(str, s)))
def test_set_equilibrium_equals_loading_feq(mods, storage, shape):
    torch, amd = mods
    print_case("set-equilibrium-one-pass", storage=storage, shape=shape)
    rng = np.random.default_rng(11)
    rho = 1.0 + 0.01 * rng.standard_normal(shape)
    u = 0.03 * rng.standard_normal((3,) + shape)
    # FP64: host arrays; FP32: device tensors (host arrays are taken as FP64)
    r32 = torch.tensor(rho, dtype=torch.float32, device="cuda")
    u32 = torch.tensor(u, dtype=torch.float32, device="cuda")
    for name, args, feq in (("f64", (rho, u), oracle.feq_field(rho, u)),
                            ("f32", (r32, u32), amd.reference.feq_field(r32, u32, precision="fp32"))):
        a = amd.Lattice(shape, 0.7, storage=storage).set_equilibrium(*args)
        b = amd.Lattice(shape, 0.7, storage=storage).set_populations(feq)
        (ra, la, na), (rb, lb, nb) = a.raw_state(), b.raw_state()
        assert np.array_equal(ra, rb) and la == lb and na == nb == 0, (storage, name)
        a.close()
        b.close()


def y_profiles(torch, ny):
    y = torch.arange(ny, device="cuda", dtype=torch.float32)
    k = 2.0 * math.pi / ny
    zero = torch.zeros_like(y)
    u = torch.stack([0.02 * torch.sin(k * y), zero, 0.01 * torch.cos(k * y)])
    F = torch.stack([1e-6 * torch.cos(k * y), zero, 2e-6 * torch.sin(k * y)])
    return u, F


def run_columns(torch, amd, shape, steps):
    nx, ny, nz = shape
    uprof, fprof = y_profiles(torch, ny)
    F = fprof[:, None, :, None].expand(3, nx, ny, nz).contiguous()
    lat = amd.Lattice(shape, 0.6, storage="f16", force=F)
    del F
    u = uprof[:, None, :, None].expand(3, nx, ny, nz).contiguous()
    lat.set_equilibrium(None, u)
    del u
    lat.step(steps)
    rho = torch.empty(shape, dtype=torch.float32, device="cuda")
    uo = torch.empty((3,) + shape, dtype=torch.float32, device="cuda")
    lat.moments_into(rho, uo)
    lat.close()
    return rho, uo


def test_lattice_past_32bit_field_offsets(mods):
    torch, amd = mods
    shape = (1400, 16, 65536)
    ncell = math.prod(shape)
    assert 3 * ncell > 2**32
    total = torch.cuda.get_device_properties(0).total_memory
    if total < 200e9:
        pytest.skip(f"needs ~180 GB of device memory, the GPU has {total / 1e9:.0f} GB")
    steps = 30
    print_case("large-lattice", shape=shape, ncell=ncell, storage="f16", steps=steps)
    rs, us = run_columns(torch, amd, (3, 16, 64), steps)
    assert float(us[2].abs().max()) > 1e-3 and float(us[0].abs().max()) > 1e-3
    rho, u = run_columns(torch, amd, shape, steps)
    col_r = rs[:1, :, :1]
    col_u = us[:, :1, :, :1]
    worst = 0.0
    for x0 in range(0, shape[0], 100):
        x1 = min(shape[0], x0 + 100)
        worst = max(worst, float((rho[x0:x1] - col_r).abs().max()), float((u[:, x0:x1] - col_u).abs().max()))
    print(f"large lattice: max |column - small lattice| = {worst:.3e}", flush=True)
    assert worst == 0.0
    del rho, u
    torch.cuda.empty_cache()
