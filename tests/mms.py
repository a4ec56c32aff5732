"""Forced Taylor-Green MMS harness, as the reference's
packages/lattice-boltzmann-d3q19/tests/test_mms_convergence.py:75-131, with the
manufactured solution of
tools/testkit/code_verification/mms/solutions/incompressible_ns_2d/solution.py:56-92.

The step functions are injected so the same harness drives the oracle and
the GPU drop-in.
"""

from __future__ import annotations

import math

import numpy as np

CS2 = 1.0 / 3.0
TAU_TARGET, NU_PHYS, T_FINAL, A_AMP, NZ = 0.65, 0.01, 0.05, 0.05, 3


def evaluate(x, y, t):
    ct = math.cos(t)
    u = np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) * ct
    v = -np.cos(2 * np.pi * x) * np.sin(2 * np.pi * y) * ct
    return u, v


def source(x, y, t, nu=NU_PHYS):
    sx, cx = np.sin(2 * np.pi * x), np.cos(2 * np.pi * x)
    sy, cy = np.sin(2 * np.pi * y), np.cos(2 * np.pi * y)
    s4x, s4y = np.sin(4 * np.pi * x), np.sin(4 * np.pi * y)
    ct, st = math.cos(t), math.sin(t)
    e = 8 * np.pi * np.pi
    su = -sx * cy * st + 2 * np.pi * s4x * ct * ct + e * nu * sx * cy * ct
    sv = cx * sy * st + 2 * np.pi * s4y * ct * ct - e * nu * cx * sy * ct
    return su, sv


def setup(N):
    dx = 1.0 / N
    dt_nat = CS2 * (TAU_TARGET - 0.5) * dx * dx / NU_PHYS
    n_steps = max(1, int(round(T_FINAL / dt_nat)))
    dt = T_FINAL / n_steps
    tau = 0.5 + NU_PHYS * dt / (CS2 * dx * dx)
    idx = (np.arange(N, dtype=np.float64) + 0.5) / N
    X, Y = np.meshgrid(idx, idx, indexing="ij")
    return dx, dt, n_steps, tau, X, Y


def force_at(X, Y, t, N, dt, dx):
    su, sv = source(X, Y, t)
    sf = dt * dt / dx
    F = np.zeros((3, N, N, NZ))
    F[0] = (A_AMP * su * sf)[:, :, None]
    F[1] = (A_AMP * sv * sf)[:, :, None]
    return F


def initial_velocity(X, Y, N, dt, dx):
    u0, v0 = evaluate(X, Y, 0.0)
    sv = dt / dx
    ul = np.zeros((3, N, N, NZ))
    ul[0] = (A_AMP * u0 * sv)[:, :, None]
    ul[1] = (A_AMP * v0 * sv)[:, :, None]
    return ul


def error_of(u_lat_end, X, Y, dt, dx):
    sv = dt / dx
    ue = (u_lat_end[0] / sv)[:, :, 0]
    ve = (u_lat_end[1] / sv)[:, :, 0]
    ua, va = evaluate(X, Y, T_FINAL)
    eu = math.sqrt(np.mean((ue - A_AMP * ua) ** 2))
    ev = math.sqrt(np.mean((ve - A_AMP * va) ** 2))
    return math.sqrt(eu * eu + ev * ev)


def mms_error(N, bgk_step, feq_field, macroscopic_velocity):
    dx, dt, n_steps, tau, X, Y = setup(N)
    f = feq_field(np.ones((N, N, NZ)), initial_velocity(X, Y, N, dt, dx))
    for step in range(n_steps):
        f = bgk_step(f, tau, force_at(X, Y, (step + 0.5) * dt, N, dt, dx))
    u = macroscopic_velocity(f, force_at(X, Y, T_FINAL, N, dt, dx))
    return n_steps, tau, error_of(np.asarray(u), X, Y, dt, dx)
