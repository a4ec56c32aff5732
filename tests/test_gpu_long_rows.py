"""GPU checks of the fused step's two layouts and of its bit-level determinism.

The collision is written with contraction off and explicit fused
multiply-adds, so the result of a cell does not depend on the kernel it is
inlined into.  Half-way walls with no wall faces are the periodic step by
another kernel instantiation (a separate pull routine), so the periodic and
the faceless half-way lattices must agree bit for bit, in every storage and
with or without forcing.  Shapes cover rows of 128 to 1024 cells, several
blocks per row (384 = 3 x 128, 1024 = 4 x 256) and ny not a multiple of the
rows per block; and short rows (nz < 64: direction-major storage on a linear
grid, e.g. the reference's depth-3 slabs), where the last block is partial.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import field_scaled_err, print_case

pytestmark = pytest.mark.gpu

SHAPES = [(3, 5, 256), (2, 3, 512), (4, 6, 128), (3, 2, 384), (2, 7, 1024),
          (16, 8, 3), (7, 5, 9), (33, 17, 1), (5, 3, 63)]


@pytest.fixture(scope="module")
def amd():
    import lattice_boltzmann_d3q19_amd as m

    return m


def seeded(shape, seed):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.01 * rng.standard_normal(shape)
    u = 0.02 * rng.standard_normal((3,) + shape)
    f = oracle.feq_field(rho, u) * (1.0 + 0.02 * rng.standard_normal((19,) + shape))
    F = 1e-4 * rng.standard_normal((3,) + shape)
    return f, F


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16", "bf16sr"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("forcing", ["none", "field", "uniform-brinkman"])
def test_periodic_equals_faceless_halfway_bitwise(amd, storage, shape, forcing):
    import torch

    print_case("periodic-vs-halfway", storage=storage, shape=shape, forcing=forcing)
    f0, F = seeded(shape, seed=sum(shape))
    kw = {}
    if forcing == "field":
        kw["force"] = F
    elif forcing == "uniform-brinkman":
        kw["force"] = (2e-5, -1e-5, 0.0)
    outs = []
    for boundary in ("periodic", "halfway"):
        lat = amd.Lattice(shape, 0.6, storage=storage, boundary=boundary, **kw)
        lat.set_populations(f0)
        if forcing == "uniform-brinkman":
            lat.set_brinkman(np.random.default_rng(1).random(shape) * 0.2)
        lat.step(12)
        rho = torch.empty(shape, dtype=torch.float32, device="cuda")
        u = torch.empty((3,) + shape, dtype=torch.float32, device="cuda")
        lat.step_moments(3, rho, u)
        torch.cuda.synchronize()
        outs.append((lat.populations(np.float32), rho.cpu().numpy(), u.cpu().numpy()))
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("shape", [(4, 6, 256), (3, 5, 128), (2, 4, 512)], ids=lambda s: "x".join(map(str, s)))
def test_long_rows_ywalls_match_oracle(amd, shape):
    print_case("long-rows-ywalls", shape=shape)
    f0, F = seeded(shape, seed=7)
    top, bot = (0.04, 0.0, -0.01), (0.0, 0.0, 0.02)
    lat = amd.Lattice(shape, 0.65, storage="f32", boundary="ywalls",
                      wall_velocity_top=top, wall_velocity_bottom=bot, force=F)
    lat.set_populations(f0).step(40)
    ref = oracle.run_reference(f0, 0.65, 40, force_lattice=F, ywalls=True,
                               wall_velocity_top=top, wall_velocity_bottom=bot)
    assert field_scaled_err(lat.populations(), ref) < 1e-6


@pytest.mark.parametrize("shape", [(3, 4, 256), (2, 3, 384)], ids=lambda s: "x".join(map(str, s)))
def test_long_rows_periodic_match_oracle(amd, shape):
    print_case("long-rows-periodic", shape=shape)
    f0, F = seeded(shape, seed=9)
    lat = amd.Lattice(shape, 0.6, storage="f32", force=F)
    lat.set_populations(f0).step(40)
    ref = oracle.run_reference(f0, 0.6, 40, force_lattice=F)
    assert field_scaled_err(lat.populations(), ref) < 1e-6


def test_short_rows_store_no_padding(amd):
    """A depth-3 slab holds two direction-major lattices of exactly 19 * ncell
    populations; a 64-long row needs no padding either."""
    print_case("short-rows-footprint")
    for shape, es in (((256, 256, 3), 8), ((64, 32, 3), 4), ((8, 8, 64), 4)):
        storage = "f64" if es == 8 else "f32"
        lat = amd.Lattice(shape, 0.7, storage=storage)
        ncell = int(np.prod(shape))
        pops = 2 * 19 * ncell * es
        assert pops <= lat.device_bytes < pops + 64 * 1024 + ncell * 8, (shape, lat.device_bytes)
        lat.close()
