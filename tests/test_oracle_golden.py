"""Pin the FP64 oracle's equilibrium and constants against the reference's
golden table and known answers (CPU).

Reference tests mirrored: packages/lattice-boltzmann-d3q19/tests/test_d3q19_equilibrium_golden.py:29-51,
packages/lattice-boltzmann-d3q19-stack-e/tests/test_reference_sanity.py:30-60,
packages/lattice-boltzmann-d3q19/lattice_boltzmann_d3q19/invariants.py:50-131 (PBT).
"""

from __future__ import annotations

import json

import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st

import oracle
from conftest import GOLDEN


@pytest.fixture(scope="module")
def golden():
    with (GOLDEN / "d3q19-equilibrium.json").open() as fh:
        return json.load(fh)


def test_constants_match_reference_definition():
    c, w, opp, cs2 = oracle.constants()
    # lexicographic order of NP/reference/constants.py:39-59
    assert c[0].tolist() == [0, 0, 0]
    assert c[1:7].tolist() == [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]]
    assert c[7].tolist() == [1, 1, 0] and c[18].tolist() == [0, -1, 1]
    assert float(np.sum(w)) == 1.0  # exact (stack-e test_reference_sanity)
    assert cs2 == 1.0 / 3.0
    # opposite map is an involution with c[opp] = -c
    assert np.array_equal(opp[opp], np.arange(19))
    assert np.array_equal(c[opp], -c)
    # second-moment isotropy sum w c_a c_b = cs2 delta_ab
    m2 = np.einsum("i,ia,ib->ab", w, c.astype(float), c.astype(float))
    assert np.allclose(m2, cs2 * np.eye(3), atol=1e-16)


def test_feq_rest_equals_weights_exactly():
    _, w, _, _ = oracle.constants()
    assert oracle.feq(1.0, (0.0, 0.0, 0.0)) == w.tolist()


@pytest.mark.parametrize("point", range(4))
def test_feq_golden_table(golden, point):
    tp = golden["test_points"][point]
    f = oracle.feq(tp["inputs"]["rho"], tp["inputs"]["u"])
    for i, expected in enumerate(tp["expected"]["f_eq"]):
        assert f[i] == pytest.approx(expected, abs=1e-15)
    assert sum(f) == pytest.approx(tp["expected"]["density_moment"], abs=1e-14)


def test_feq_field_matches_golden_and_momentum_known_answer(golden):
    tp = golden["test_points"][0]
    rho = np.ones((1, 1, 1))
    u = np.asarray(tp["inputs"]["u"], np.float64).reshape(3, 1, 1, 1)
    f = oracle.feq_field(rho, u)
    assert np.allclose(f.reshape(19), tp["expected"]["f_eq"], atol=1e-15, rtol=0)
    # momentum_moment of the point feq (equilibrium.py:34-68), as the golden test does
    fp = np.asarray(oracle.feq(tp["inputs"]["rho"], tp["inputs"]["u"])).reshape(19, 1, 1, 1)
    mom = oracle.momentum_field(fp).reshape(3)
    # the reference golden test compares at abs 1e-14 (test_d3q19_equilibrium_golden.py:48-51)
    assert mom[0] == pytest.approx(tp["expected"]["momentum_x"], abs=1e-14)
    assert mom[1] == pytest.approx(0.0, abs=1e-14) and mom[2] == pytest.approx(0.0, abs=1e-14)
    assert oracle.density_field(f).reshape(()) == pytest.approx(1.0, abs=1e-14)


_U = 0.1 * (1.0 / 3.0) ** 0.5  # |u| < 0.1 c_s band of invariants.py


@settings(max_examples=50, deadline=None)
@given(rho=st.floats(0.5, 1.5), ux=st.floats(-_U, _U), uy=st.floats(-_U, _U), uz=st.floats(-_U, _U))
def test_pbt_equilibrium_moments(rho, ux, uy, uz):
    f = oracle.feq(rho, (ux, uy, uz))
    assert abs(sum(f) - rho) <= 1e-14 * max(abs(rho), 1.0)
    c, _, _, _ = oracle.constants()
    mom = (np.asarray(f)[:, None] * c).sum(axis=0)
    for a, ua in enumerate((ux, uy, uz)):
        assert abs(mom[a] - rho * ua) <= 1e-14 * max(abs(rho * ua), 1.0)


def test_field_moments_brute_force():
    rng = np.random.default_rng(7)
    f = rng.random((19, 3, 4, 5))
    c, _, _, _ = oracle.constants()
    rho = oracle.density_field(f)
    mom = oracle.momentum_field(f)
    assert np.allclose(rho, f.sum(axis=0), rtol=1e-15, atol=0)
    assert np.allclose(mom, np.einsum("ia,ixyz->axyz", c.astype(float), f), rtol=1e-14, atol=1e-15)
    F = rng.standard_normal((3, 3, 4, 5)) * 1e-3
    u = oracle.macroscopic_velocity(f, F)
    assert np.allclose(u, (mom + 0.5 * F) / rho, rtol=1e-14, atol=0)
