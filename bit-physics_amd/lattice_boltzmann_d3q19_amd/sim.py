"""Canonical runs of the reference, stepped on the GPU and captured to disk.

Same runners, descriptors, parameters, step order and diagnostics as
packages/lattice-boltzmann-d3q19/lattice_boltzmann_d3q19/sim.py:

* :func:`sim_runner_seeded` -- Poiseuille 64 x 32 x 3, tau 0.7, F_x = 1e-5,
  reference y-walls, 1000 steps, every step captured (sim.py:288-336,429-475);
* :func:`sim_runner_seeded_couette` -- Couette 32 x 16 x 3, top wall
  u = (0.05, 0, 0), 500 steps (sim.py:339-376,478-520);
* :func:`sim_runner_diagnostic` -- Poiseuille 16 x 8 x 3, 50 steps, every 10th
  captured (sim.py:523-549);
* :func:`sim_runner_moment_encoded` -- the canonical Poiseuille on 16-bit
  moment-encoded storage, as the frontier-moment-encoded variant's capture
  (packages/lattice-boltzmann-d3q19-frontier-moment-encoded/src/lbm_me_capture_main.cpp).

Each capture frame holds rho = density_field(f^n) and
u = macroscopic_velocity(f^n, F) in float64, plus the tier-1 health and
tier-2 vector-field diagnostics (sim.py:135-162).  The state advances with
the engine's fused step (``Lattice.step``) between frames; rho and u come
from the engine's read-back.  ``storage="f64"`` (the default, the
reference's precision) or ``"f32"``.
"""

from __future__ import annotations

import datetime
import json
import time
from pathlib import Path
from typing import Iterator

import numpy as np

from .capture import CaptureManifest, StepState, write_capture
from .lattice import Lattice
from .reference import (
    CANONICAL_COUETTE_NX,
    CANONICAL_COUETTE_NY,
    CANONICAL_COUETTE_STEPS,
    CANONICAL_DESCRIPTOR_COUETTE,
    CANONICAL_DESCRIPTOR_POISEUILLE,
    CANONICAL_NZ,
    CANONICAL_POISEUILLE_NX,
    CANONICAL_POISEUILLE_NY,
    CANONICAL_POISEUILLE_STEPS,
)

CANONICAL_TAU = 0.7
CANONICAL_POISEUILLE_FORCE_X = 1.0e-5
CANONICAL_COUETTE_WALL_VELOCITY = 0.05
DIAGNOSTIC_DESCRIPTOR = "poiseuille-16x8-seed42-step50"

STACK = {"name": "mi355x-hip", "version": "0.1.0", "build_id": "libd3q19-gfx950"}


def health_check(rho: np.ndarray, u: np.ndarray) -> dict:
    """Tier-1 health (sim.py:135-145): 0.0 when every value is finite, else
    1.0; density range; largest |u|."""
    finite = bool(np.all(np.isfinite(rho)) and np.all(np.isfinite(u)))
    return {
        "check_health": 0.0 if finite else 1.0,
        "rho_min": float(np.min(rho)),
        "rho_max": float(np.max(rho)),
        "u_max_lat": float(np.max(np.sqrt((u * u).sum(axis=0)))),
    }


def vector_field_diagnostics(u: np.ndarray, dx_phys: float) -> dict:
    """Tier-2 vector-field diagnostics on the z = 0 slice (sim.py:148-162):
    max |div u| by periodic central differences in x and y, and the
    circulation sum(u_x) dx^2."""
    ux, uy = u[0, :, :, 0], u[1, :, :, 0]
    h = 0.5 / dx_phys
    div = (np.roll(ux, -1, axis=0) - np.roll(ux, 1, axis=0)) * h + \
          (np.roll(uy, -1, axis=1) - np.roll(uy, 1, axis=1)) * h
    return {
        "check_divergence_free": float(np.max(np.abs(div))),
        "check_circulation": float(np.sum(ux) * dx_phys * dx_phys),
    }


def _frame(step: int, rho: np.ndarray, u: np.ndarray, dx_phys: float) -> StepState:
    return StepState(step=step,
                     state={"rho": rho.astype(np.float64), "u": u.astype(np.float64)},
                     diagnostics={**health_check(rho, u), **vector_field_diagnostics(u, dx_phys)})


def evolve(lattice: Lattice, n_steps: int, capture_interval: int, dx_phys: float) -> Iterator[StepState]:
    """Frames of a run from the lattice's loaded state: step 0, then every
    ``capture_interval``-th step and the last one (sim.py:306-326)."""
    rho, u = lattice.moments()
    yield _frame(0, rho, u, dx_phys)
    done = 0
    while done < n_steps:
        nxt = min(n_steps, (done // capture_interval + 1) * capture_interval)
        lattice.step(nxt - done)
        done = nxt
        rho, u = lattice.moments()
        yield _frame(done, rho, u, dx_phys)


def poiseuille_lattice(nx, ny, nz, tau=CANONICAL_TAU, force_x=CANONICAL_POISEUILLE_FORCE_X,
                       storage="f64") -> Lattice:
    """Rest state (rho = 1, u = 0) between reference y-walls, uniform body
    force F = (force_x, 0, 0) (sim.py:281-285,298-300)."""
    return Lattice((nx, ny, nz), tau, storage=storage, boundary="ywalls",
                   force=(force_x, 0.0, 0.0)).set_equilibrium()


def couette_lattice(nx, ny, nz, tau=CANONICAL_TAU, wall_velocity=CANONICAL_COUETTE_WALL_VELOCITY,
                    storage="f64") -> Lattice:
    """Rest state, top wall moving at (wall_velocity, 0, 0) (sim.py:350-364)."""
    return Lattice((nx, ny, nz), tau, storage=storage, boundary="ywalls",
                   wall_velocity_top=(wall_velocity, 0.0, 0.0)).set_equilibrium()


def _manifest(descriptor, tier, dims, seed, params, n_steps, interval, storage) -> CaptureManifest:
    return CaptureManifest(
        schema_version="1.0.0",
        sim={"name": "lattice-boltzmann-d3q19", "category": "lattice", "variant": "bgk-d3q19-qian-1992"},
        stack=dict(STACK),
        config={"tier": tier, "dims": list(dims), "dtype": "f64" if storage == "f64" else "f32",
                "seed": int(seed), "params": params},
        run={"step_count": int(n_steps), "capture_interval": int(interval), "wall_clock_seconds": 0.0,
             "start_utc": datetime.datetime.now(datetime.timezone.utc).strftime("%Y-%m-%dT%H:%M:%SZ")},
        payload={"format": "hdf5", "path": descriptor + ".h5", "checksum": "sha256:" + "0" * 64},
        determinism={"claimed": "bit-exact-same-hw", "atomic_ops": False, "subgroup_ops": False},
    )


def _run(lattice, manifest, n_steps, interval, dx_phys, out_dir, payload_format) -> Path:
    t0 = time.perf_counter()
    path = write_capture(evolve(lattice, n_steps, interval, dx_phys), manifest, out_dir, payload_format)
    lattice.close()
    m = json.loads(path.read_text())
    m["run"]["wall_clock_seconds"] = time.perf_counter() - t0
    path.write_text(json.dumps(m, indent=2, sort_keys=True))
    return path


def sim_runner_seeded(seed: int, out_dir, storage: str = "f64", payload_format=None) -> Path:
    """Canonical Poiseuille capture; returns the manifest path (sim.py:429-475)."""
    nx, ny, nz = CANONICAL_POISEUILLE_NX, CANONICAL_POISEUILLE_NY, CANONICAL_NZ
    params = {"descriptor": CANONICAL_DESCRIPTOR_POISEUILLE, "tau": CANONICAL_TAU,
              "force_x_lattice": CANONICAL_POISEUILLE_FORCE_X, "nz_convention": "depth-3-z-periodic-slab",
              "boundary": "bounce-back-y-walls-periodic-xz", "storage": storage}
    m = _manifest(CANONICAL_DESCRIPTOR_POISEUILLE, "test", (nx, ny, nz), seed, params,
                  CANONICAL_POISEUILLE_STEPS, 1, storage)
    return _run(poiseuille_lattice(nx, ny, nz, storage=storage), m, CANONICAL_POISEUILLE_STEPS, 1,
                1.0 / nx, out_dir, payload_format)


def sim_runner_seeded_couette(seed: int, out_dir, storage: str = "f64", payload_format=None) -> Path:
    """Canonical Couette capture (sim.py:478-520)."""
    nx, ny, nz = CANONICAL_COUETTE_NX, CANONICAL_COUETTE_NY, CANONICAL_NZ
    params = {"descriptor": CANONICAL_DESCRIPTOR_COUETTE, "tau": CANONICAL_TAU,
              "wall_velocity_top_lattice": CANONICAL_COUETTE_WALL_VELOCITY,
              "nz_convention": "depth-3-z-periodic-slab",
              "boundary": "bounce-back-y-walls-with-moving-top-plate-periodic-xz", "storage": storage}
    m = _manifest(CANONICAL_DESCRIPTOR_COUETTE, "test", (nx, ny, nz), seed, params,
                  CANONICAL_COUETTE_STEPS, 1, storage)
    return _run(couette_lattice(nx, ny, nz, storage=storage), m, CANONICAL_COUETTE_STEPS, 1,
                1.0 / nx, out_dir, payload_format)


def sim_runner_diagnostic(seed: int, out_dir, storage: str = "f64", payload_format=None) -> Path:
    """Diagnostic-tier Poiseuille 16 x 8 x 3, 50 steps, frames every 10 (sim.py:523-549)."""
    nx, ny, nz, n, interval = 16, 8, CANONICAL_NZ, 50, 10
    params = {"descriptor": DIAGNOSTIC_DESCRIPTOR, "tau": CANONICAL_TAU, "force_x_lattice": 1.0e-5,
              "nz_convention": "depth-3-z-periodic-slab", "boundary": "bounce-back-y-walls-periodic-xz",
              "storage": storage}
    m = _manifest(DIAGNOSTIC_DESCRIPTOR, "diagnostic", (nx, ny, nz), seed, params, n, interval, storage)
    return _run(poiseuille_lattice(nx, ny, nz, force_x=1.0e-5, storage=storage), m, n, interval,
                1.0 / nx, out_dir, payload_format)


def moment_encoded_diagnostics(rho: np.ndarray, u: np.ndarray) -> dict:
    """Frame diagnostics of the moment-encoded variant's capture
    (ME/src/lbm_me.cpp:465-474): density range and the largest |u| component."""
    return {"rho_min": float(np.min(rho)), "rho_max": float(np.max(rho)), "u_max_lat": float(np.max(np.abs(u)))}


def sim_runner_moment_encoded(seed: int, out_dir, steps: int = CANONICAL_POISEUILLE_STEPS,
                              payload_format=None) -> Path:
    """Canonical Poiseuille on 16-bit moment-encoded storage, every step
    captured (ME/src/lbm_me_capture_main.cpp, lbm_me.cpp:395-486): ranges
    calibrated over the whole horizon (margin 0.25), frame 0 is the encoded
    initial state.  The codes are stepped in the variant's own FP64
    arithmetic ("me16ref"), so the frames are the variant's."""
    from .moment_encoded import RANGE_MARGIN, WARMUP_STEPS, moment_encoded_lattice

    nx, ny, nz = CANONICAL_POISEUILLE_NX, CANONICAL_POISEUILLE_NY, CANONICAL_NZ
    t0 = time.perf_counter()
    lat = moment_encoded_lattice((nx, ny, nz), CANONICAL_TAU, steps, boundary="ywalls",
                                 force=(CANONICAL_POISEUILLE_FORCE_X, 0.0, 0.0), reference_arithmetic=True)
    params = {"tau": CANONICAL_TAU, "force_x_lattice": CANONICAL_POISEUILLE_FORCE_X, "quantize_bits": 16.0,
              "range_margin": RANGE_MARGIN, "warmup_steps": float(WARMUP_STEPS)}
    m = _manifest(CANONICAL_DESCRIPTOR_POISEUILLE, "test", (nx, ny, nz), seed, params, steps, 1, "f64")
    m.sim["variant"] = "frontier-moment-encoded-16bit"

    def frames():
        for n in range(steps + 1):
            if n:
                lat.step(1)
            rho, u = lat.moments()
            yield StepState(step=n, state={"rho": rho, "u": u}, diagnostics=moment_encoded_diagnostics(rho, u))

    path = write_capture(frames(), m, out_dir, payload_format)
    lat.close()
    mm = json.loads(path.read_text())
    mm["run"]["wall_clock_seconds"] = time.perf_counter() - t0
    path.write_text(json.dumps(mm, indent=2, sort_keys=True))
    return path


__all__ = [
    "CANONICAL_COUETTE_WALL_VELOCITY", "CANONICAL_POISEUILLE_FORCE_X", "CANONICAL_TAU",
    "DIAGNOSTIC_DESCRIPTOR", "couette_lattice", "evolve", "health_check", "poiseuille_lattice",
    "moment_encoded_diagnostics", "sim_runner_diagnostic", "sim_runner_moment_encoded", "sim_runner_seeded",
    "sim_runner_seeded_couette", "vector_field_diagnostics",
]
