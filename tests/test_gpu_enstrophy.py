"""FP64 enstrophy reduction (d3q19_enstrophy) against a brute-force NumPy curl
and the discrete closed form of the Taylor-Green field (no reference
counterpart: a diagnostic of the decaying-turbulence workload, whose
comparison covers kinetic energy, enstrophy and total mass).

Enstrophy = 0.5 sum |curl u|^2 with periodic central differences
0.5 (u(x+1) - u(x-1)), u the read-back (Guo-corrected) velocity.  For the
sampled Taylor-Green field u = (u0 sin X cos Y cos Z, -u0 cos X sin Y cos Z, 0),
X = 2 pi x / nx, ..., a central difference of sin(k x) is sin(k) cos(k x), so
  curl u = u0 (-s_z cos X sin Y sin Z, -s_z sin X cos Y sin Z, (s_x + s_y) sin X sin Y cos Z),
  s_a = sin(2 pi / n_a), and with sum sin^2 = sum cos^2 = n / 2 along each axis
  enstrophy = 0.5 u0^2 (ncell / 8) (2 s_z^2 + (s_x + s_y)^2).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from conftest import print_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import lattice_boltzmann_d3q19_amd as amd

    return amd


def curl_enstrophy(u):
    """0.5 sum |curl u|^2, periodic central differences, in NumPy (FP64)."""
    u = np.asarray(u, np.float64)

    def d(a, axis):
        return 0.5 * (np.roll(a, -1, axis) - np.roll(a, 1, axis))

    ux, uy, uz = u
    ox = d(uz, 1) - d(uy, 2)
    oy = d(ux, 2) - d(uz, 0)
    oz = d(uy, 0) - d(ux, 1)
    return 0.5 * float(np.sum(ox * ox + oy * oy + oz * oz))


def taylor_green_enstrophy(shape, u0):
    s = [math.sin(2 * math.pi / n) for n in shape]
    return 0.5 * u0 * u0 * (math.prod(shape) / 8.0) * (2 * s[2] ** 2 + (s[0] + s[1]) ** 2)


@pytest.mark.parametrize("storage,tol", [("f64", 1e-12), ("f32", 1e-12), ("f16", 1e-12)])
def test_enstrophy_matches_brute_force_curl(amd, storage, tol):
    """A seeded 24x20x16 lattice with a force field, 5 steps: the reduction of
    the read-back u equals the NumPy curl of the same u."""
    print_case("enstrophy-brute-force", storage=storage, shape=(24, 20, 16))
    shape = (24, 20, 16)
    rng = np.random.default_rng(42)
    f0 = oracle.feq_field(1.0 + 0.01 * rng.standard_normal(shape), 0.03 * rng.standard_normal((3,) + shape))
    F = 1e-4 * rng.standard_normal((3,) + shape)
    lat = amd.Lattice(shape, 0.7, storage=storage, force=F).set_populations(f0).step(5)
    _, u = lat.moments()
    got = lat.enstrophy()
    lat.close()
    want = curl_enstrophy(u)
    assert abs(got - want) <= tol * want, (got, want)


@pytest.mark.parametrize("shape", [(16, 12, 20), (32, 32, 32)])
def test_enstrophy_of_taylor_green_is_the_discrete_closed_form(amd, shape):
    from lattice_boltzmann_d3q19_amd.initial import taylor_green

    print_case("enstrophy-taylor-green", shape=shape)
    u0 = 0.04
    rho, u = taylor_green(shape, u0)
    lat = amd.Lattice(shape, 0.6, storage="f64").set_equilibrium(rho, u)
    got = lat.enstrophy()
    lat.close()
    want = taylor_green_enstrophy(shape, u0)
    assert abs(got - want) <= 1e-12 * want, (got, want)
    # the continuous derivative (k in place of sin k) is a different number
    k = [2 * math.pi / n for n in shape]
    continuous = 0.5 * u0 * u0 * (math.prod(shape) / 8.0) * (2 * k[2] ** 2 + (k[0] + k[1]) ** 2)
    assert abs(continuous - want) > 1e-3 * want


def test_enstrophy_refuses_lattices_with_walls_or_solids(amd):
    from lattice_boltzmann_d3q19_amd import D3Q19Error

    print_case("enstrophy-refuses-walls")
    shape = (8, 6, 4)
    cases = [
        dict(boundary="ywalls"),
        dict(boundary="halfway", wall_faces=("-z", "+z")),
    ]
    for kw in cases:
        lat = amd.Lattice(shape, 0.6, storage="f32", **kw).set_equilibrium()
        with pytest.raises(D3Q19Error, match="EINVAL"):
            lat.enstrophy()
        lat.close()
    solid = np.zeros(shape, np.uint8)
    solid[2:4, 2:4, 1:3] = 1
    lat = amd.Lattice(shape, 0.6, storage="f32", boundary="halfway").set_solid(solid).set_equilibrium()
    with pytest.raises(D3Q19Error, match="EINVAL"):
        lat.enstrophy()
    lat.close()
    # half-way mode without walls or solids is periodic: same value as the periodic lattice
    from lattice_boltzmann_d3q19_amd.initial import taylor_green

    rho, u = taylor_green(shape, 0.03)
    a = amd.Lattice(shape, 0.6, storage="f64", boundary="halfway").set_equilibrium(rho, u)
    b = amd.Lattice(shape, 0.6, storage="f64").set_equilibrium(rho, u)
    assert a.enstrophy() == b.enstrophy() == pytest.approx(taylor_green_enstrophy(shape, 0.03), rel=1e-12)
    a.close()
    b.close()
