"""Pin the FP64 oracle's step against reference known answers, golden
vectors, closed forms and finite differences (CPU).

Known answers:
  * Poiseuille 64x32x3, 1000 steps: max|u| 8.6508320e-03, max rho 1.0000000;
    Couette 32x16x3, 500 steps: max|u| 5.0000000e-02
    (docs/sim-specs/lattice/lattice-boltzmann-d3q19/equivalence.md:238-243).
  * MMS ladder errors 4.224e-3 / 3.033e-3 / 1.543e-4 at N = 32/64/128, OOA 2.39
    (docs/_audits/phase-1/sub-phase-lattice-boltzmann-d3q19/landing-2026-05-23T00-41-15Z.md:178-183),
    harness as packages/lattice-boltzmann-d3q19/tests/test_mms_convergence.py:75-131.
  * Rest state is a fixed point, mass/momentum conserved without force
    (packages/lattice-boltzmann-d3q19-frontier-moment-encoded/tests/test_lbm_me.cpp:96-113).
Golden vectors: tests/golden/ywalls_moving.npz (written by the NumPy reference).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from mms import mms_error


def rest(nx, ny, nz):
    return oracle.feq_field(np.ones((nx, ny, nz)), np.zeros((3, nx, ny, nz)))


def uniform(F, shape):
    return np.broadcast_to(np.asarray(F, np.float64)[:, None, None, None], (3,) + shape).copy()


def test_poiseuille_known_answer():
    nx, ny, nz = 64, 32, 3
    F = uniform((1e-5, 0, 0), (nx, ny, nz))
    f = oracle.run_reference(rest(nx, ny, nz), 0.7, 1000, force_lattice=F, ywalls=True)
    u = oracle.macroscopic_velocity(f, F)
    rho = oracle.density_field(f)
    assert f"{np.max(np.sqrt((u * u).sum(0))):.7e}" == "8.6508320e-03"
    assert f"{rho.max():.7e}" == "1.0000000e+00"


def test_couette_known_answer():
    nx, ny, nz = 32, 16, 3
    f = oracle.run_reference(rest(nx, ny, nz), 0.7, 500, ywalls=True,
                             wall_velocity_top=(0.05, 0.0, 0.0))
    u = oracle.macroscopic_velocity(f)
    assert f"{np.max(np.sqrt((u * u).sum(0))):.7e}" == "5.0000000e-02"


def test_ywalls_golden_vectors_bit_exact():
    d = np.load(GOLDEN / "ywalls_moving.npz", allow_pickle=False)
    bb = oracle.apply_bounce_back_y_walls(d["f0"], d["wall_top"], d["wall_bottom"])
    assert np.array_equal(bb, d["bb_once"])
    f = d["f0"]
    for k in range(d["composite"].shape[0]):
        f = oracle.bgk_step(f, float(d["tau"]), d["force"])
        f = oracle.apply_bounce_back_y_walls(f, d["wall_top"], d["wall_bottom"])
        assert np.array_equal(f, d["composite"][k]), f"composite step {k + 1}"


def test_ywalls_golden_catches_planted_sign_error():
    """A flipped moving-wall term must fail the golden comparison."""
    d = np.load(GOLDEN / "ywalls_moving.npz", allow_pickle=False)
    wrong = oracle.apply_bounce_back_y_walls(d["f0"], -d["wall_top"], -d["wall_bottom"])
    assert not np.allclose(wrong, d["bb_once"], rtol=0, atol=1e-12)


def test_mms_ladder_known_answers():
    errs = [mms_error(N, oracle.bgk_step, oracle.feq_field, oracle.macroscopic_velocity)[2]
            for N in (32, 64, 128)]
    assert [f"{e:.3e}" for e in errs] == ["4.224e-03", "3.033e-03", "1.543e-04"]
    slope = float(np.polyfit(np.log([1 / 32, 1 / 64, 1 / 128]), np.log(errs), 1)[0])
    assert f"{slope:.2f}" == "2.39"


def test_rest_state_is_fixed_point():
    f0 = rest(6, 5, 4)
    f = oracle.run_reference(f0, 0.7, 3, ywalls=True)
    assert np.max(np.abs(f - f0)) <= 1e-15
    f = oracle.run_halfway(f0, 0.7, 3, wall_mask=63)
    assert np.max(np.abs(f - f0)) <= 1e-15


def test_periodic_conservation_without_force():
    rng = np.random.default_rng(42)
    shape = (8, 6, 5)
    f = oracle.feq_field(1 + 0.01 * rng.standard_normal(shape), 0.01 * rng.standard_normal((3,) + shape))
    d0 = oracle.diagnostics(f)
    for _ in range(10):
        f = oracle.bgk_step(f, 0.8)
    d1 = oracle.diagnostics(f)
    assert abs(d1[0] - d0[0]) <= 1e-12 * abs(d0[0])
    assert np.all(np.abs(d1[1:4] - d0[1:4]) <= 1e-12)


def test_uniform_force_adds_momentum_exactly():
    shape = (5, 4, 3)
    f = rest(*shape)
    F = (1e-4, -2e-4, 3e-4)
    g = oracle.bgk_step(f, 0.7, uniform(F, shape))
    d = oracle.diagnostics(g)
    n = np.prod(shape)
    assert np.allclose(d[1:4], np.asarray(F) * n, rtol=1e-12, atol=0)


def test_stream_is_np_roll():
    rng = np.random.default_rng(3)
    f = rng.random((19, 5, 4, 3))
    c, _, _, _ = oracle.constants()
    expect = np.stack([np.roll(f[i], shift=tuple(c[i]), axis=(0, 1, 2)) for i in range(19)])
    assert np.array_equal(oracle.stream(f), expect)


def test_halfway_poiseuille_exact_at_magic_tau():
    """Half-way bounce-back + Guo: parabola exact at tau = 1/2 + sqrt(3/16)."""
    nx, ny, nz = 2, 12, 2
    tau = 0.5 + math.sqrt(3.0 / 16.0)
    nu = (tau - 0.5) / 3.0
    g = 1e-6
    f = oracle.run_halfway(rest(nx, ny, nz), tau, 20000, wall_mask=(1 << 2) | (1 << 3),
                           force_uniform=(g, 0, 0))
    u = oracle.macroscopic_velocity(f, uniform((g, 0, 0), (nx, ny, nz)))
    y = np.arange(ny)
    ua = g / (2 * nu) * (y + 0.5) * (ny - 0.5 - y)
    assert np.max(np.abs(u[0, 0, :, 0] - ua)) / ua.max() < 1e-9


def test_halfway_couette_linear_profile():
    nx, ny, nz = 2, 10, 2
    U = 0.01
    fv = np.zeros((6, 3))
    fv[3] = (U, 0, 0)
    f = oracle.run_halfway(rest(nx, ny, nz), 0.8, 12000, wall_mask=(1 << 2) | (1 << 3), face_velocity=fv)
    u = oracle.macroscopic_velocity(f)
    ua = U * (np.arange(ny) + 0.5) / ny
    assert np.max(np.abs(u[0, 0, :, 0] - ua)) / U < 1e-9


def test_halfway_without_walls_equals_periodic_step():
    rng = np.random.default_rng(5)
    shape = (4, 5, 3)
    f = oracle.feq_field(1 + 0.01 * rng.standard_normal(shape), 0.01 * rng.standard_normal((3,) + shape))
    F = 1e-4 * rng.standard_normal((3,) + shape)
    a = oracle.run_halfway(f, 0.7, 2, force_field=F)
    b = oracle.bgk_step(oracle.bgk_step(f, 0.7, F), 0.7, F)
    assert np.max(np.abs(a - b)) < 1e-16


def test_halfway_cavity_conserves_mass():
    shape = (6, 6, 6)
    fv = np.zeros((6, 3))
    fv[5] = (0.05, 0, 0)
    f0 = rest(*shape)
    f = oracle.run_halfway(f0, 0.6, 50, wall_mask=63, face_velocity=fv)
    d0, d1 = oracle.diagnostics(f0), oracle.diagnostics(f)
    assert abs(d1[0] - d0[0]) < 1e-12 * d0[0]
    u = oracle.macroscopic_velocity(f)
    assert u[0, :, :, -1].mean() > 0  # the lid drags the top layer in +x


def test_brinkman_full_solid_fraction_removes_momentum():
    rng = np.random.default_rng(9)
    shape = (3, 3, 3)
    f = oracle.feq_field(np.ones(shape), 0.02 * rng.standard_normal((3,) + shape))
    s = np.ones(shape)
    g = oracle.run_halfway(f, 1.0, 1, brinkman=s)
    # s = 1: post-collision momentum is zero at every cell, streaming keeps the total at zero
    assert np.all(np.abs(oracle.diagnostics(g)[1:4]) < 1e-15)


def test_inverse_adjoint_matches_central_differences():
    rng = np.random.default_rng(42)
    shape = (4, 3, 5)
    tau, T = 0.6, 6
    u0 = 0.02 * rng.standard_normal((3,) + shape)
    F = 1e-4 * rng.standard_normal((3,) + shape)
    ut = 0.02 * rng.standard_normal((3,) + shape)
    L, gu, gF = oracle.inverse_grad(u0, F, ut, tau, T)
    assert L == pytest.approx(oracle.inverse_loss(u0, F, ut, tau, T), rel=1e-14)
    for k in range(4):
        du = rng.standard_normal(u0.shape)
        dF = rng.standard_normal(F.shape)
        eps = 1e-6
        fd = (oracle.inverse_loss(u0 + eps * du, F + eps * dF, ut, tau, T)
              - oracle.inverse_loss(u0 - eps * du, F - eps * dF, ut, tau, T)) / (2 * eps)
        ad = float(np.sum(gu * du) + np.sum(gF * dF))
        assert ad == pytest.approx(fd, rel=1e-6), k
