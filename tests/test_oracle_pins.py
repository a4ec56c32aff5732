"""Independent pins of the oracle's additions that no reference vector covers
(CPU): the FP64 diagnostics (mass, momentum, kinetic energy) and the solid
cells of the half-way streaming.

Each check is written from the definition or the closed form, never from the
oracle's own formula, and each is shown to have power: a test build of the
oracle with one planted mistake (``oracle.planted``) must fail it.

* Diagnostics: mass = sum f, momentum = sum c f, kinetic energy
  KE = 0.5 sum rho |u|^2 with the Guo-corrected velocity
  u = (sum c f + F / 2) / max(rho, 1e-30)
  (packages/lattice-boltzmann-d3q19/lattice_boltzmann_d3q19/reference/bgk.py:129-142),
  summed over fluid cells only when a solid mask is given.
* Solid cells (half-way bounce-back from a solid source): the fluid mass of a
  box of stationary solid blocks is conserved; a channel between two solid
  rows y = 0 and y = ny - 1 has its walls half way, at y = 1/2 and
  y = ny - 3/2, so at tau = 1/2 + sqrt(3/16) the Guo-forced flow is the
  parabola u(y) = g (y - 1/2)(ny - 3/2 - y) / (2 nu) to rounding
  (Krueger et al. 2017, sec. 5.3.3: exact for this tau).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle

# Lexicographic D3Q19 velocity set of the reference
# (packages/lattice-boltzmann-d3q19/lattice_boltzmann_d3q19/reference/constants.py:39-60),
# written out here so the checks do not read the oracle's table.
C = np.array([
    (0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
    (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1),
    (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1)], dtype=np.float64)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def seeded_populations(shape, seed, amp_u=0.05, noneq=1e-3):
    """Equilibrium of a perturbed (rho, u) plus a non-equilibrium perturbation."""
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.05 * rng.standard_normal(shape)
    u = amp_u * rng.standard_normal((3,) + shape)
    f = oracle.feq_field(rho, u)
    return f + noneq * W[:, None, None, None] * rng.standard_normal((19,) + shape)


def diagnostics_from_definition(f, force=None, solid=None):
    """[mass, px, py, pz, KE] from the definitions, in NumPy."""
    rho = f.sum(axis=0)
    m = np.tensordot(C.T, f, axes=(1, 0))
    F = np.zeros(3) if force is None else np.asarray(force, np.float64)
    u = (m + 0.5 * F[:, None, None, None]) / np.maximum(rho, 1e-30)
    ke = 0.5 * rho * (u * u).sum(axis=0)
    fluid = np.ones(rho.shape, bool) if solid is None else ~np.asarray(solid, bool)
    return np.array([rho[fluid].sum(), m[0][fluid].sum(), m[1][fluid].sum(), m[2][fluid].sum(), ke[fluid].sum()])


CASES = [
    ("unforced", None, False),
    ("uniform force", (2e-3, -1e-3, 3e-3), False),
    ("force and solids", (2e-3, -1e-3, 3e-3), True),
]


def _diag_case(force, with_solid, seed=42):
    shape = (7, 6, 5)
    f = seeded_populations(shape, seed)
    solid = None
    if with_solid:
        solid = np.random.default_rng(seed + 1).random(shape) < 0.2
    got = oracle.diagnostics(f, force_uniform=force, solid=solid)
    want = diagnostics_from_definition(f, force, solid)
    return got, want


def _rel(got, want):
    return np.abs(got - want) / np.abs(want)


@pytest.mark.parametrize("name,force,with_solid", CASES, ids=[c[0] for c in CASES])
def test_diagnostics_match_definitions(name, force, with_solid):
    got, want = _diag_case(force, with_solid)
    assert _rel(got[0], want[0]) <= 1e-13
    # momentum sums are O(1) with cancellations: compare against the sum of |c f|
    assert np.all(np.abs(got[1:4] - want[1:4]) <= 1e-13 * 210)
    assert _rel(got[4], want[4]) <= 1e-13, (got[4], want[4])


@pytest.mark.parametrize("plant", ["KE_NO_HALF", "KE_UNIT_RHO", "KE_NO_GUO"])
def test_kinetic_energy_check_rejects_planted_mistakes(plant):
    with oracle.planted(plant):
        got, want = _diag_case((2e-3, -1e-3, 3e-3), True)
    assert _rel(got[4], want[4]) > 1e-6, plant
    assert _rel(got[0], want[0]) <= 1e-13  # mass is untouched by these plants
    got, want = _diag_case((2e-3, -1e-3, 3e-3), True)  # the correct build is back
    assert _rel(got[4], want[4]) <= 1e-13


def rest(shape):
    return oracle.feq_field(np.ones(shape), np.zeros((3,) + shape))


def solid_blocks(shape):
    s = np.zeros(shape, np.uint8)
    s[2:5, 1:4, 2:6] = 1        # a block
    s[8:10, 5:9, 0:3] = 1       # a block touching a periodic face
    s[0, :, 7] = 1              # a thin wall across the periodic x seam
    return s


def _solid_box_mass_drift(steps=500):
    shape = (12, 10, 8)
    solid = solid_blocks(shape)
    fluid = solid == 0
    rng = np.random.default_rng(7)
    f0 = oracle.feq_field(1.0 + 0.01 * rng.standard_normal(shape), 0.02 * rng.standard_normal((3,) + shape))
    fv = np.zeros((6, 3))
    f = oracle.run_halfway(f0, 0.6, steps, solid=solid, force_uniform=(1e-5, 2e-5, 0.0))
    m0 = f0[:, fluid].sum()
    m1 = f[:, fluid].sum()
    del fv
    return abs(m1 - m0) / m0, f, solid


def test_solid_blocks_conserve_fluid_mass():
    drift, f, solid = _solid_box_mass_drift()
    assert drift <= 1e-13, drift
    # solid cells hold the rest weights after every streaming
    assert np.array_equal(f[:, solid.astype(bool)], np.repeat(W[:, None], int(solid.sum()), axis=1))


@pytest.mark.parametrize("plant", ["SOLID_SAME_DIR", "SOLID_WRONG_SIDE"])
def test_solid_mass_check_rejects_planted_mistakes(plant):
    with oracle.planted(plant):
        drift, _, _ = _solid_box_mass_drift(steps=50)
    assert drift > 1e-8, (plant, drift)


def _solid_row_channel(ny=14, steps=20000):
    nx, nz = 2, 2
    tau = 0.5 + math.sqrt(3.0 / 16.0)
    nu = (tau - 0.5) / 3.0
    g = 1e-6
    solid = np.zeros((nx, ny, nz), np.uint8)
    solid[:, 0, :] = 1
    solid[:, ny - 1, :] = 1
    f = oracle.run_halfway(rest((nx, ny, nz)), tau, steps, solid=solid, force_uniform=(g, 0, 0))
    u = oracle.macroscopic_velocity(f, np.array([g, 0, 0])[:, None, None, None] * np.ones((3, nx, ny, nz)))
    y = np.arange(1, ny - 1)
    ua = g / (2 * nu) * (y - 0.5) * (ny - 1.5 - y)
    return f, u[0, :, 1:ny - 1, :], ua, tau, g


def test_channel_between_solid_rows_is_the_half_way_parabola():
    f, ux, ua, tau, g = _solid_row_channel()
    assert np.max(np.abs(ux - ua[None, :, None])) / ua.max() < 1e-9
    # the same channel bounded by wall faces (ny - 2 fluid rows) runs the same arithmetic
    fw = oracle.run_halfway(rest((2, 12, 2)), tau, 20000, wall_mask=(1 << 2) | (1 << 3), force_uniform=(g, 0, 0))
    assert np.array_equal(f[:, :, 1:-1, :], fw)


@pytest.mark.parametrize("plant", ["SOLID_SAME_DIR", "SOLID_WRONG_SIDE"])
def test_channel_check_rejects_planted_mistakes(plant):
    with oracle.planted(plant):
        _, ux, ua, _, _ = _solid_row_channel(steps=4000)
    _, ux_ok, _, _, _ = _solid_row_channel(steps=4000)
    # after 4000 steps the correct flow is still developing: compare the two runs
    assert np.max(np.abs(ux - ux_ok)) / ua.max() > 1e-3, plant
