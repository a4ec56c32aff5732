"""The manufactured solution of the MMS ladder, pinned to the PDE it solves.

The source must equal the residual of the incompressible Navier-Stokes
momentum equation for the solution's velocity and its pressure
p = -(cos 4 pi x + cos 4 pi y) cos^2 t / 4
(tools/testkit/code_verification/mms/solutions/incompressible_ns_2d/solution.py:56-66):
    S = u_t + (u . grad) u + grad p - nu lap u,
checked here by central differences, and the velocity must be divergence-free.
The ladder's known answers pin the rest (tests/test_oracle_step.py).
"""

from __future__ import annotations

import numpy as np

from lattice_boltzmann_d3q19_amd import verification as V


def pressure(x, y, t):
    return -0.25 * (np.cos(4 * np.pi * x) + np.cos(4 * np.pi * y)) * np.cos(t) ** 2


def test_source_is_navier_stokes_residual():
    rng = np.random.default_rng(0)
    x, y = rng.random(40), rng.random(40)
    h, nu = 1e-4, V.NU_PHYS
    for t in (0.0, 0.02, 0.05, 0.7):
        u, v = V.evaluate(x, y, t)
        ut = (np.stack(V.evaluate(x, y, t + h)) - np.stack(V.evaluate(x, y, t - h))) / (2 * h)
        dx = (np.stack(V.evaluate(x + h, y, t)) - np.stack(V.evaluate(x - h, y, t))) / (2 * h)
        dy = (np.stack(V.evaluate(x, y + h, t)) - np.stack(V.evaluate(x, y - h, t))) / (2 * h)
        lap = (np.stack(V.evaluate(x + h, y, t)) + np.stack(V.evaluate(x - h, y, t)) +
               np.stack(V.evaluate(x, y + h, t)) + np.stack(V.evaluate(x, y - h, t)) -
               4 * np.stack(V.evaluate(x, y, t))) / (h * h)
        px = (pressure(x + h, y, t) - pressure(x - h, y, t)) / (2 * h)
        py = (pressure(x, y + h, t) - pressure(x, y - h, t)) / (2 * h)
        res = ut + u * dx + v * dy + np.stack([px, py]) - nu * lap
        su, sv = V.source(x, y, t)
        assert np.max(np.abs(res - np.stack([su, sv]))) < 1e-5, t
        assert np.max(np.abs(dx[0] + dy[1])) < 1e-7


def test_ladder_levels():
    # diffusive scaling: tau stays near 0.65 and n_steps grows as N^2
    for N in (32, 64, 128, 1024):
        dx, dt, n, tau, X, Y = V.setup(N)
        assert abs(tau - V.TAU_TARGET) < 2e-2 and abs(n * dt - V.T_FINAL) < 1e-15
    assert V.setup(32)[2] == 10 and V.setup(128)[2] == 164 and V.setup(1024)[2] == 10486
    assert abs(V.observed_order([32, 64, 128], [4.224e-3, 3.033e-3, 1.543e-4]) - 2.39) < 5e-3
