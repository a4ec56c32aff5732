"""Half-way z walls on the rows' fast path.

A row off the x and y wall faces, in a lattice without solid cells, takes the
periodic pull and then bounces its z-wall lanes (for 16-bit storage the wall
lanes load their bounce values before the pull).  A lattice given an all-fluid
solid mask takes the general per-direction path instead, so the two must
agree bit for bit: every storage, moving walls on both z faces (alone and
with x/y walls), and Nz = 1 (a lane on both z walls), 2, 3 (short rows) and
64, 70 (row-blocked layout).  The 16-bit runs are also held to the FP64
oracle at the storage-matrix tolerances.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import print_case

pytestmark = pytest.mark.gpu

FACES = {"z": ("-z", "+z"), "all": ("-x", "+x", "-y", "+y", "-z", "+z")}
FV = {"-z": (0.02, -0.01, 0.0), "+z": (-0.03, 0.02, 0.0), "+y": (0.04, 0.0, 0.01)}
SHAPES = [(5, 4, 1), (5, 4, 2), (4, 3, 3), (4, 3, 64), (3, 4, 70)]
TOL = {"f16": 1e-3, "bf16": 1.2e-2, "bf16sr": 2e-2}


@pytest.fixture(scope="module")
def amd():
    import lattice_boltzmann_d3q19_amd as amd

    return amd


def run(amd, shape, storage, faces, u, steps, solid_mask):
    lat = amd.Lattice(shape, 0.62, storage=storage, boundary="halfway", wall_faces=FACES[faces],
                      face_velocity={k: v for k, v in FV.items() if k in FACES[faces]})
    if solid_mask:
        lat.set_solid(np.zeros(shape, np.uint8))
    lat.set_equilibrium(None, u).step(steps)
    f = lat.populations()
    lat.close()
    return f


@pytest.mark.parametrize("storage", ["f64", "f32", "f16", "bf16", "bf16sr"])
@pytest.mark.parametrize("faces", list(FACES))
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_fast_path_equals_general_path(amd, storage, faces, shape):
    print_case("halfway-zwalls-fast-vs-general", storage=storage, faces=faces, shape=shape)
    u = 0.02 * np.random.default_rng(7).standard_normal((3,) + shape)
    fast = run(amd, shape, storage, faces, u, 30, False)
    general = run(amd, shape, storage, faces, u, 30, True)
    assert np.isfinite(fast).all()
    assert np.array_equal(fast, general)


@pytest.mark.parametrize("storage", list(TOL))
@pytest.mark.parametrize("shape", [(5, 4, 2), (4, 3, 64)], ids=lambda s: "x".join(map(str, s)))
def test_16bit_zwalls_match_oracle(amd, storage, shape):
    print_case("halfway-zwalls-oracle", storage=storage, shape=shape)
    u = 0.02 * np.random.default_rng(8).standard_normal((3,) + shape)
    f = run(amd, shape, storage, "z", u, 60, False)
    fv = np.zeros((6, 3))
    fv[4], fv[5] = FV["-z"], FV["+z"]
    ref = oracle.run_halfway(oracle.feq_field(np.ones(shape), u), 0.62, 60, wall_mask=48, face_velocity=fv)
    _, w, _, _ = oracle.constants()
    d, dref = f - w[:, None, None, None], ref - w[:, None, None, None]
    assert np.abs(d - dref).max() / np.abs(dref).max() < TOL[storage]
