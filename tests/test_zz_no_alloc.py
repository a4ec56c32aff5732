"""Allocation-free moment read-back (d3q19_read_moments_noalloc) through the C
ABI, against the allocating d3q19_read_moments on the same lattice: bit for
bit, on two calls in a row (the second on a smaller lattice, so stale staged
values of the first would show), for rho and u together and apart, and the
size limit of its fixed staging buffer.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle
from conftest import print_case

pytestmark = pytest.mark.gpu

LIMIT = 16 * 1024 * 1024  # D3Q19_NOALLOC_SCRATCH_BYTES


@pytest.fixture(scope="module")
def amd():
    import lattice_boltzmann_d3q19_amd as amd

    return amd


def _p(a):
    return ctypes.c_void_p(None if a is None else a.ctypes.data)


def read_both(lat, dtype, parts=("rho", "u")):
    from lattice_boltzmann_d3q19_amd._lib import D3Q19_F32, D3Q19_F64, call

    code = D3Q19_F64 if dtype == np.float64 else D3Q19_F32
    shape = lat.shape
    ref = {"rho": np.empty(shape, dtype), "u": np.empty((3,) + shape, dtype)}
    call("d3q19_read_moments", lat._h, _p(ref["rho"]), _p(ref["u"]), code, 0)
    got = {k: (np.full(v.shape, np.nan, dtype) if k in parts else None) for k, v in ref.items()}
    call("d3q19_read_moments_noalloc", lat._h, _p(got["rho"]), _p(got["u"]), code)
    return ref, got


def seeded(shape, seed):
    rng = np.random.default_rng(seed)
    return oracle.feq_field(1.0 + 0.01 * rng.standard_normal(shape), 0.03 * rng.standard_normal((3,) + shape))


def test_noalloc_read_back_equals_read_moments_on_two_calls(amd):
    # first call: FP64 storage with a per-cell force field, read as FP64
    print_case("noalloc-first", shape=(24, 20, 16), storage="f64")
    shape = (24, 20, 16)
    F = 1e-4 * np.random.default_rng(3).standard_normal((3,) + shape)
    lat = amd.Lattice(shape, 0.7, storage="f64", force=F).set_populations(seeded(shape, 1)).step(7)
    ref, got = read_both(lat, np.float64)
    assert np.array_equal(ref["rho"], got["rho"]) and np.array_equal(ref["u"], got["u"])
    lat.close()
    # second call, smaller: an FP32 half-way cavity with a moving lid, read as FP32
    print_case("noalloc-second", shape=(12, 10, 8), storage="f32")
    shape = (12, 10, 8)
    lat = amd.Lattice(shape, 0.6, storage="f32", boundary="halfway",
                      wall_faces=("-x", "+x", "-y", "+y", "-z", "+z"),
                      face_velocity={"+z": (0.05, 0.0, 0.0)}).set_equilibrium().step(30)
    ref, got = read_both(lat, np.float32)
    assert np.array_equal(ref["rho"], got["rho"]) and np.array_equal(ref["u"], got["u"])
    # either output alone
    for part in ("rho", "u"):
        ref, got = read_both(lat, np.float32, parts=(part,))
        assert np.array_equal(ref[part], got[part])
    lat.close()


def test_noalloc_read_back_size_limit(amd):
    from lattice_boltzmann_d3q19_amd import D3Q19Error
    from lattice_boltzmann_d3q19_amd._lib import D3Q19_F64, call

    shape = (128, 64, 65)  # 532,480 cells: 32 B per cell in FP64 is over the limit, 16 B in FP32 is not
    assert 4 * 8 * np.prod(shape) > LIMIT >= 4 * 4 * np.prod(shape)
    print_case("noalloc-limit", shape=shape)
    lat = amd.Lattice(shape, 0.6, storage="f32").set_populations(seeded(shape, 2))
    rho = np.empty(shape, np.float64)
    with pytest.raises(D3Q19Error, match="EINVAL"):
        call("d3q19_read_moments_noalloc", lat._h, _p(rho), _p(None), D3Q19_F64)
    ref, got = read_both(lat, np.float32)
    assert np.array_equal(ref["rho"], got["rho"]) and np.array_equal(ref["u"], got["u"])
    lat.close()
