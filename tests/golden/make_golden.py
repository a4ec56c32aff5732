"""Write reference input/output vectors for the moving-wall bounce-back.

Run in the build container only (the reference tree is not present where
the GPU tests run):  python tests/golden/make_golden.py

Only the function that no closed form, published known answer or
reference-held fixture pins to the bit is exercised here:
``apply_bounce_back_y_walls`` with MOVING walls, whose momentum term uses a
wall density summed over post-stream populations that include the ones
wrapped periodically from the opposite wall
(packages/lattice-boltzmann-d3q19/lattice_boltzmann_d3q19/reference/bgk.py:145-216),
alone and composed with ``bgk_step`` + Guo force as in NP/sim.py:312-326.

Output: tests/golden/ywalls_moving.npz (float64 arrays; load with
numpy.load(..., allow_pickle=False)).
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

REF_PKG = Path("/root/reference/packages/lattice-boltzmann-d3q19")


def main() -> None:
    sys.path.insert(0, str(REF_PKG))
    from lattice_boltzmann_d3q19.reference import (  # noqa: E402
        apply_bounce_back_y_walls,
        bgk_step,
        feq_field,
    )

    nx, ny, nz = 8, 6, 4
    tau = 0.7
    rng = np.random.default_rng(42)
    rho0 = 1.0 + 0.01 * rng.standard_normal((nx, ny, nz))
    u0 = 0.01 * rng.standard_normal((3, nx, ny, nz))
    f0 = feq_field(rho0, u0)
    # non-equilibrium part so every population differs
    f0 = f0 * (1.0 + 0.02 * rng.standard_normal(f0.shape))
    force = 1e-4 * rng.standard_normal((3, nx, ny, nz))
    top = (0.05, -0.01, 0.02)
    bot = (-0.03, 0.0, 0.01)

    bb_once = apply_bounce_back_y_walls(f0, wall_velocity_top=top, wall_velocity_bottom=bot)

    f = f0.copy()
    composite = []
    for _ in range(3):
        f = bgk_step(f, tau, force_lattice=force)
        f = apply_bounce_back_y_walls(f, wall_velocity_top=top, wall_velocity_bottom=bot)
        composite.append(f.copy())

    out = Path(__file__).resolve().parent / "ywalls_moving.npz"
    np.savez(
        out,
        f0=f0,
        force=force,
        tau=np.float64(tau),
        wall_top=np.asarray(top, np.float64),
        wall_bottom=np.asarray(bot, np.float64),
        bb_once=bb_once,
        composite=np.stack(composite),
    )
    print(f"wrote {out} ({out.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
