"""Initial fields for the engine's workloads (NumPy float64, host side).

* :func:`taylor_green` -- the 3-D Taylor-Green vortex on a 2*pi periodic box
  with its consistent pressure:
      u = (u0 sin x cos y cos z, -u0 cos x sin y cos z, 0),
      p = (u0^2 / 16)(cos 2x + cos 2y)(cos 2z + 2),  rho = 1 + p / c_s^2,
  x_i = 2 pi i / N.  At low Reynolds number the kinetic energy decays as
  KE(t) = KE(0) exp(-2 nu |k|^2 t) with |k|^2 = 3 (2 pi / N)^2.
* :func:`random_solenoidal` -- a seeded random-phase, divergence-free
  velocity field with energy spectrum E(k) ~ k^4 exp(-2 (k / k0)^2),
  scaled to a target rms velocity.

The reference has no such generators (its scenes start from rest,
packages/lattice-boltzmann-d3q19/lattice_boltzmann_d3q19/sim.py:281-285).
"""

from __future__ import annotations

import numpy as np


def taylor_green(shape, u0: float = 0.04):
    nx, ny, nz = shape
    x = 2 * np.pi * np.arange(nx) / nx
    y = 2 * np.pi * np.arange(ny) / ny
    z = 2 * np.pi * np.arange(nz) / nz
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    u = np.empty((3, nx, ny, nz), np.float64)
    u[0] = u0 * np.sin(X) * np.cos(Y) * np.cos(Z)
    u[1] = -u0 * np.cos(X) * np.sin(Y) * np.cos(Z)
    u[2] = 0.0
    p = (u0 * u0 / 16.0) * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
    rho = 1.0 + 3.0 * p
    return rho, u


def taylor_green_ke_decay(shape, nu: float, t: float) -> float:
    """KE(t)/KE(0) of the linear (low-Re) Taylor-Green solution."""
    nx, ny, nz = shape
    k2 = (2 * np.pi / nx) ** 2 + (2 * np.pi / ny) ** 2 + (2 * np.pi / nz) ** 2
    return float(np.exp(-2.0 * nu * k2 * t))


def random_solenoidal(shape, u_rms: float, seed: int = 42, k0: float = 8.0, dtype=np.float64):
    """Seeded random-phase solenoidal field with E(k) ~ k^4 exp(-2 (k/k0)^2)."""
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(nx, d=1.0 / nx)[:, None, None]
    ky = np.fft.fftfreq(ny, d=1.0 / ny)[None, :, None]
    kz = np.fft.rfftfreq(nz, d=1.0 / nz)[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    k = np.sqrt(k2)
    # white noise has uniformly random phases; shape its amplitude so that the
    # shell-summed energy follows E(k): per-mode |u_k|^2 ~ E(k) / k^2
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where(k > 0, np.sqrt(k ** 4 * np.exp(-2.0 * (k / k0) ** 2) / np.maximum(k2, 1e-30)), 0.0)
    # Nyquist modes of even axes carry no velocity: their wavenumber has no sign
    # (+-n/2 alias), so the inverse real transform would undo their projection.
    # (Where the spectrum has decayed there, e.g. 384^3 with k0 = 8, they are 0.)
    nyq = np.zeros(amp.shape, bool)
    if nx % 2 == 0:
        nyq |= np.abs(kx) == nx // 2
    if ny % 2 == 0:
        nyq |= np.abs(ky) == ny // 2
    if nz % 2 == 0:
        nyq |= kz == nz // 2
    amp = np.where(nyq, 0.0, amp)
    uh = []
    for _ in range(3):
        w = rng.standard_normal(shape)
        uh.append(np.fft.rfftn(w) * amp)
    # Leray projection: remove the component along k
    with np.errstate(divide="ignore", invalid="ignore"):
        kdotu = (kx * uh[0] + ky * uh[1] + kz * uh[2]) / np.where(k2 > 0, k2, 1.0)
    uh = [uh[0] - kx * kdotu, uh[1] - ky * kdotu, uh[2] - kz * kdotu]
    u = np.stack([np.fft.irfftn(c, s=shape, axes=(0, 1, 2)) for c in uh]).astype(np.float64)
    rms = np.sqrt(np.mean(np.sum(u * u, axis=0)) / 3.0)
    u *= u_rms / rms
    return u.astype(dtype)


def spectral_divergence(u) -> float:
    """max |k . u_k| / max |k| |u_k| -- zero for a solenoidal field."""
    shape = u.shape[1:]
    kx = np.fft.fftfreq(shape[0], d=1.0 / shape[0])[:, None, None]
    ky = np.fft.fftfreq(shape[1], d=1.0 / shape[1])[None, :, None]
    kz = np.fft.rfftfreq(shape[2], d=1.0 / shape[2])[None, None, :]
    uh = [np.fft.rfftn(c) for c in u]
    div = kx * uh[0] + ky * uh[1] + kz * uh[2]
    mag = np.sqrt(kx * kx + ky * ky + kz * kz) * np.sqrt(sum(np.abs(c) ** 2 for c in uh))
    return float(np.max(np.abs(div)) / max(np.max(mag), 1e-300))


def random_spheres(shape, r_min=6.0, r_max=12.0, volume_fraction=0.05, seed=42, wall_axis=1):
    """Seeded random spherical obstacles as a Brinkman solid fraction in [0, 1].

    Spheres of radius r in [r_min, r_max] are placed at uniform random centres
    (periodic along the axes other than ``wall_axis``; kept clear of the two
    walls along ``wall_axis``) until the solid volume reaches
    ``volume_fraction``.  The fraction falls linearly from 1 to 0 over one
    cell at the sphere surface: s = clip(r + 0.5 - distance, 0, 1).
    """
    rng = np.random.default_rng(seed)
    n = np.asarray(shape)
    s = np.zeros(shape, np.float64)
    grids = np.meshgrid(*[np.arange(k, dtype=np.float64) for k in shape], indexing="ij")
    target = volume_fraction * s.size
    for _ in range(100000):
        if s.sum() >= target:
            break
        r = rng.uniform(r_min, r_max)
        c = rng.uniform(0, n)
        c[wall_axis] = rng.uniform(r + 1, n[wall_axis] - r - 2)
        d2 = np.zeros(shape)
        for a in range(3):
            d = grids[a] - c[a]
            if a != wall_axis:
                d = d - n[a] * np.round(d / n[a])  # periodic distance
            d2 += d * d
        s = np.maximum(s, np.clip(r + 0.5 - np.sqrt(d2), 0.0, 1.0))
    return s
