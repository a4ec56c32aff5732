"""Solution verification: the forced Taylor-Green manufactured solution.

The reference's MMS ladder
(packages/lattice-boltzmann-d3q19/tests/test_mms_convergence.py:75-165) runs
the periodic BGK+Guo step on an N x N x 3 slab with the incompressible 2-D
Navier-Stokes solution of
tools/testkit/code_verification/mms/solutions/incompressible_ns_2d/solution.py:56-92,

    u = A sin(2 pi x) cos(2 pi y) cos t,   v = -A cos(2 pi x) sin(2 pi y) cos t,

driven by its source term S(x, y, t) at the step midpoint, diffusive scaling
dt ~ dx^2 (tau ~ 0.65, nu = 0.01, A = 0.05, t_final = 0.05), and reports the
RMS velocity error at t_final.  Its known answers are 4.224e-3, 3.033e-3,
1.543e-4 at N = 32, 64, 128 (observed order 2.39).

* :func:`mms_error` drives any functional ``bgk_step`` / ``feq_field`` /
  ``macroscopic_velocity`` triple (the reference's loop);
* :func:`mms_error_lattice` drives the engine's fused step: the source is
  S = sin t B1 + cos^2 t B2 + cos t B3 with three fixed device fields, so each
  step costs three AXPYs, a device copy into the lattice's force field and
  one fused step, and the ladder reaches N = 1024 and beyond.
"""

from __future__ import annotations

import math
import time

import numpy as np

CS2 = 1.0 / 3.0
TAU_TARGET, NU_PHYS, T_FINAL, A_AMP, NZ = 0.65, 0.01, 0.05, 0.05, 3
FORMAL_ORDER = 2.0
KNOWN_ANSWERS = {32: 4.224e-3, 64: 3.033e-3, 128: 1.543e-4}


def evaluate(x, y, t):
    """(u, v) of the unit-amplitude solution at time t."""
    ct = math.cos(t)
    u = np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) * ct
    v = -np.cos(2 * np.pi * x) * np.sin(2 * np.pi * y) * ct
    return u, v


def _basis(x, y, nu=NU_PHYS):
    """S(x, y, t) = sin t B1 + cos^2 t B2 + cos t B3, each B a (u, v) pair."""
    sx, cx = np.sin(2 * np.pi * x), np.cos(2 * np.pi * x)
    sy, cy = np.sin(2 * np.pi * y), np.cos(2 * np.pi * y)
    e = 8 * np.pi * np.pi * nu
    b1 = (-sx * cy, cx * sy)
    b2 = (2 * np.pi * np.sin(4 * np.pi * x), 2 * np.pi * np.sin(4 * np.pi * y))
    b3 = (e * sx * cy, -e * cx * sy)
    return b1, b2, b3


def source(x, y, t, nu=NU_PHYS):
    """Momentum source of the unit-amplitude solution (u_t + u.grad u - nu lap u)."""
    b1, b2, b3 = _basis(x, y, nu)
    st, ct = math.sin(t), math.cos(t)
    return tuple(st * b1[k] + ct * ct * b2[k] + ct * b3[k] for k in range(2))


def setup(N):
    """dx, dt, number of steps, tau and the cell-centred grid of level N."""
    dx = 1.0 / N
    dt_nat = CS2 * (TAU_TARGET - 0.5) * dx * dx / NU_PHYS
    n_steps = max(1, int(round(T_FINAL / dt_nat)))
    dt = T_FINAL / n_steps
    tau = 0.5 + NU_PHYS * dt / (CS2 * dx * dx)
    idx = (np.arange(N, dtype=np.float64) + 0.5) / N
    X, Y = np.meshgrid(idx, idx, indexing="ij")
    return dx, dt, n_steps, tau, X, Y


def force_at(X, Y, t, N, dt, dx):
    """Lattice force (3, N, N, 3): A S dt^2 / dx."""
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


def mach_max(u_lat) -> float:
    return float(np.max(np.sqrt((u_lat * u_lat).sum(axis=0)))) / math.sqrt(CS2)


def error_of(u_lat_end, X, Y, dt, dx):
    """(err_u, err_v, err): RMS errors of the physical velocity at t_final."""
    sv = dt / dx
    ue = (np.asarray(u_lat_end[0]) / sv)[:, :, 0]
    ve = (np.asarray(u_lat_end[1]) / sv)[:, :, 0]
    ua, va = evaluate(X, Y, T_FINAL)
    eu = math.sqrt(np.mean((ue - A_AMP * ua) ** 2))
    ev = math.sqrt(np.mean((ve - A_AMP * va) ** 2))
    return eu, ev, math.sqrt(eu * eu + ev * ev)


def mms_error(N, bgk_step, feq_field, macroscopic_velocity):
    """(n_steps, tau, err) of level N through functional step routines."""
    dx, dt, n_steps, tau, X, Y = setup(N)
    f = feq_field(np.ones((N, N, NZ)), initial_velocity(X, Y, N, dt, dx))
    for step in range(n_steps):
        f = bgk_step(f, tau, force_at(X, Y, (step + 0.5) * dt, N, dt, dx))
    u = macroscopic_velocity(f, force_at(X, Y, T_FINAL, N, dt, dx))
    return n_steps, tau, error_of(u, X, Y, dt, dx)[2]


def pressure(x, y, t):
    """Pressure of the unit-amplitude solution (solution.py:64)."""
    return -0.25 * (np.cos(4 * np.pi * x) + np.cos(4 * np.pi * y)) * math.cos(t) ** 2


def initial_density(X, Y, N, dt, dx):
    """rho = 1 + p_lat / c_s^2 with the pressure of the forced flow in lattice
    units, p_lat = p (dt / dx)^2.  The ladder scales the whole unit-amplitude
    source by A, including its gradient part u.grad u + grad p = 2 grad p
    (for this solution u.grad u = grad p), so the velocity A u solves the
    forced equations with the pressure (2 A - A^2) p."""
    p_lat = (2 * A_AMP - A_AMP * A_AMP) * pressure(X, Y, 0.0) * (dt / dx) ** 2
    return np.repeat((1.0 + p_lat / CS2)[:, :, None], NZ, axis=2)


def consistent_populations(X, Y, N, dt, dx, tau):
    """f = feq(rho, u) + f1 at t = 0: the solution's density (pressure) and the
    first-order Chapman-Enskog non-equilibrium part
        f1_i = -tau w_i rho / c_s^2 (c_i c_i - c_s^2 I) : grad u
    (lattice units), so that the state carries the viscous stress of the
    solution from the first step instead of building it up."""
    from .reference import VELOCITIES, WEIGHTS

    c = np.asarray(VELOCITIES, dtype=np.float64)
    w = np.asarray(WEIGHTS, dtype=np.float64)
    rho = initial_density(X, Y, N, dt, dx)[:, :, 0]
    u0, v0 = evaluate(X, Y, 0.0)
    s = A_AMP * dt / dx
    ux, uy = s * u0, s * v0
    # lattice-unit gradients of (ux, uy): d/dx_lat = dx d/dx
    tp = 2 * np.pi * dx
    sx, cx = np.sin(2 * np.pi * X), np.cos(2 * np.pi * X)
    sy, cy = np.sin(2 * np.pi * Y), np.cos(2 * np.pi * Y)
    g = np.zeros((3, 3) + X.shape)  # g[a, b] = d u_b / d x_a
    g[0, 0] = s * tp * cx * cy
    g[1, 0] = -s * tp * sx * sy
    g[0, 1] = s * tp * sx * sy
    g[1, 1] = -s * tp * cx * cy
    f = np.empty((19,) + X.shape)
    for i in range(19):
        cu = c[i, 0] * ux + c[i, 1] * uy
        usq = ux * ux + uy * uy
        feq = w[i] * rho * (1 + cu / CS2 + cu * cu / (2 * CS2 * CS2) - usq / (2 * CS2))
        qg = sum((c[i, a] * c[i, b] - (CS2 if a == b else 0.0)) * g[a, b] for a in range(3) for b in range(3))
        f[i] = feq - tau * w[i] * rho / CS2 * qg
    return np.repeat(f[..., None], NZ, axis=3)


def mms_error_lattice(N, storage="f64", device=0, consistent_pressure=False, consistent_init=False):
    """Level N on the engine's fused step; returns a dict of the run.

    The lattice holds a per-cell force field that is rewritten on the device
    before every step with the midpoint source; the read-back velocity uses
    the source at t_final, as the reference's loop does.  The reference
    starts from rho = 1; ``consistent_pressure`` starts from the density of
    the solution's pressure instead.
    """
    from . import _device as dev
    from .lattice import Lattice

    torch = dev.torch()
    dx, dt, n_steps, tau, X, Y = setup(N)
    u0 = initial_velocity(X, Y, N, dt, dx)
    ma = mach_max(u0)
    if ma >= 0.1:
        raise RuntimeError(f"Mach {ma:.4f} >= 0.1 at N={N}")
    ct = torch.float64 if storage == "f64" else torch.float32
    dvc = torch.device("cuda", device)
    sf = A_AMP * dt * dt / dx

    def field(pair):
        b = np.zeros((3, N, N, NZ))
        b[0] = (sf * pair[0])[:, :, None]
        b[1] = (sf * pair[1])[:, :, None]
        return torch.from_numpy(b).to(device=dvc, dtype=ct)

    b1, b2, b3 = (field(p) for p in _basis(X, Y))
    F = torch.zeros((3, N, N, NZ), device=dvc, dtype=ct)

    def set_source(t):
        s, c = math.sin(t), math.cos(t)
        torch.mul(b1, s, out=F)
        F.add_(b2, alpha=c * c)
        F.add_(b3, alpha=c)

    lat = Lattice((N, N, NZ), tau, storage=storage, force=F, device=device)
    if consistent_init:
        lat.set_populations(consistent_populations(X, Y, N, dt, dx, tau))
    else:
        lat.set_equilibrium(rho=initial_density(X, Y, N, dt, dx) if consistent_pressure else None, u=u0)
    torch.cuda.synchronize(dvc)
    t0 = time.perf_counter()
    for step in range(n_steps):
        set_source((step + 0.5) * dt)
        lat.set_force_field(F)
        lat.step(1)
    set_source(T_FINAL)
    lat.set_force_field(F)
    _, u = lat.moments()
    seconds = time.perf_counter() - t0
    lat.close()
    eu, ev, err = error_of(u, X, Y, dt, dx)
    return {"N": N, "storage": storage, "consistent_pressure": bool(consistent_pressure),
            "consistent_init": bool(consistent_init),
            "n_steps": n_steps, "tau": tau, "ma_max": ma,
            "dx": dx, "err_u": eu, "err_v": ev, "err": err, "seconds": seconds}


def observed_order(Ns, errs) -> float:
    """Least-squares slope of log err against log dx."""
    return float(np.polyfit(np.log(1.0 / np.asarray(Ns, dtype=np.float64)), np.log(errs), 1)[0])
