"""GPU checks of the fused step launched in plane chunks.

A large lattice's step runs as several launches over runs of x planes
(d3q19_launch.hpp: step_chunks); the environment variable D3Q19_STEP_CHUNKS,
read when a lattice is created, forces the number of launches.  A cell's
result does not depend on the launch it runs in, so a lattice stepped in
chunks -- including a last, shorter chunk and more chunks than planes --
must hold the same bits as one stepped in a single launch, in every storage,
boundary mode and force mode, with the moments written in the same pass.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
from conftest import print_case

pytestmark = pytest.mark.gpu

SHAPE = (23, 10, 64)


@pytest.fixture(scope="module")
def amd():
    import lattice_boltzmann_d3q19_amd as m

    return m


def make(amd, chunks, storage, **kw):
    old = os.environ.get("D3Q19_STEP_CHUNKS")
    os.environ["D3Q19_STEP_CHUNKS"] = str(chunks)
    try:
        return amd.Lattice(SHAPE, 0.6, storage=storage, **kw)
    finally:
        if old is None:
            del os.environ["D3Q19_STEP_CHUNKS"]
        else:
            os.environ["D3Q19_STEP_CHUNKS"] = old


def seeded(seed):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.01 * rng.standard_normal(SHAPE)
    u = 0.03 * rng.standard_normal((3,) + SHAPE)
    return oracle.feq_field(rho, u) * (1.0 + 0.01 * rng.standard_normal((19,) + SHAPE))


CASES = {
    "periodic": dict(),
    "periodic-forced": dict(force=(2e-5, -1e-5, 3e-6)),
    "ywalls": dict(boundary="ywalls", force=(1e-5, 0.0, 0.0), wall_velocity_top=(0.04, 0.0, 0.01)),
    "cavity": dict(boundary="halfway", wall_faces=("-x", "+x", "-y", "+y", "-z", "+z"),
                   face_velocity={"+z": (0.05, 0.0, 0.0), "-x": (0.0, 0.02, 0.0)}),
}


@pytest.mark.parametrize("storage", ["f64", "f32", "f16", "bf16sr"])
@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("chunks", [4, 7, 100])
def test_chunked_step_is_bit_identical(amd, storage, case, chunks):
    """23 planes in 4 launches (6, 6, 6 and 5 planes), 7 asked for (launches
    of 4 planes: five of 4 and one of 3) and 100 asked for (clamped to one
    launch per plane) against one launch: raw state, read-back, moments
    written by the last step and FP64 diagnostics, bit for bit."""
    print_case("step-chunks", storage=storage, case=case, chunks=chunks, shape=SHAPE)
    import torch

    f0 = seeded(7)
    out = []
    for n in (1, chunks):
        lat = make(amd, n, storage, **CASES[case])
        rho = torch.empty(SHAPE, dtype=torch.float32, device="cuda")
        u = torch.empty((3,) + SHAPE, dtype=torch.float32, device="cuda")
        lat.set_populations(f0).step(12).step_moments(3, rho, u)
        torch.cuda.synchronize()
        out.append((lat.raw_state()[0].copy(), lat.populations(), rho.cpu().numpy(), u.cpu().numpy(),
                    lat.diagnostics()))
        lat.close()
    (s1, f1, r1, u1, d1), (s2, f2, r2, u2, d2) = out
    assert np.array_equal(s1, s2)
    assert np.array_equal(f1, f2)
    assert np.array_equal(r1, r2)
    assert np.array_equal(u1, u2)
    assert d1 == d2


def test_auto_chunks_match_one_launch_at_256_planes(amd):
    """A lattice large enough to be chunked by default (FP32, 256 x 256 x 256:
    two launches per step) against the same lattice forced to one launch."""
    from lattice_boltzmann_d3q19_amd.initial import taylor_green

    shape = (256, 256, 256)
    print_case("step-chunks-auto", storage="f32", shape=shape)
    rho, u = taylor_green(shape, 0.04)
    states = []
    for n in (1, 0):
        old = os.environ.get("D3Q19_STEP_CHUNKS")
        os.environ["D3Q19_STEP_CHUNKS"] = str(n)
        try:
            lat = amd.Lattice(shape, 0.55, storage="f32")
        finally:
            if old is None:
                del os.environ["D3Q19_STEP_CHUNKS"]
            else:
                os.environ["D3Q19_STEP_CHUNKS"] = old
        lat.set_equilibrium(rho, u).step(6)
        states.append(lat.raw_state()[0].copy())
        lat.close()
    assert np.array_equal(states[0], states[1])
