"""16-bit moment-encoded storage ("me16"): range calibration and set-up.

The reference's frontier-moment-encoded variant
(packages/lattice-boltzmann-d3q19-frontier-moment-encoded/src/lbm_me.cpp:222-253,
``calibrate_ranges``) quantizes each of the 19 moments over a range taken from
an FP64 run of the whole trajectory: the minimum and maximum of the moment
over every cell of every frame 0 .. max(warmup_steps, steps), widened on both
sides by ``range_margin`` times the span (a span below 1e-12 counts as
1e-12), so that no moment clamps.  :func:`calibrate_moment_ranges` runs that
FP64 trajectory on the GPU (an "f64" lattice of the same configuration,
``Lattice.moment_range`` per frame); :func:`moment_encoded_lattice` returns an
"me16" lattice with those ranges and the initial state loaded.
"""

from __future__ import annotations

import numpy as np

from .lattice import Lattice

RANGE_MARGIN = 0.25   # lbm_me.hpp LbmConfig::range_margin
WARMUP_STEPS = 64     # lbm_me.hpp LbmConfig::warmup_steps
MIN_SPAN = 1e-12      # lbm_me.cpp:247


def _load_state(lat: Lattice, populations=None, rho=None, u=None):
    if populations is not None:
        lat.set_populations(populations)
    else:
        lat.set_equilibrium(rho, u)
    return lat


def pad_ranges(lo, hi, margin=RANGE_MARGIN):
    """Widen [lo, hi] by margin * span on both sides (span floored at 1e-12)."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    span = np.maximum(hi - lo, MIN_SPAN)
    return lo - margin * span, hi + margin * span


def calibrate_moment_ranges(shape, tau, steps, *, margin=RANGE_MARGIN, warmup_steps=WARMUP_STEPS,
                            populations=None, rho=None, u=None, prepare=None, **lattice_kw):
    """(lo, hi) of the 19 moments over the FP64 trajectory of ``steps`` steps
    (at least ``warmup_steps``) from the given initial state, padded.

    ``lattice_kw`` are the Lattice arguments of the run (boundary, walls,
    force, ...) and ``prepare(lattice)`` sets what they do not (solids, a
    Brinkman field, a force field); the "me16" lattice must be configured the
    same way.
    """
    lat = Lattice(shape, tau, storage="f64", **lattice_kw)
    if prepare is not None:
        prepare(lat)
    lat = _load_state(lat, populations, rho, u)
    lo, hi = lat.moment_range()
    for _ in range(max(int(warmup_steps), int(steps))):
        lat.step(1)
        a, b = lat.moment_range()
        lo = np.minimum(lo, a)
        hi = np.maximum(hi, b)
    lat.close()
    return pad_ranges(lo, hi, margin)


def moment_encoded_lattice(shape, tau, steps, *, ranges=None, margin=RANGE_MARGIN, warmup_steps=WARMUP_STEPS,
                           populations=None, rho=None, u=None, prepare=None, reference_arithmetic=False,
                           **lattice_kw):
    """An "me16" lattice with calibrated (or given) ranges and the initial
    state loaded; ``steps`` is the horizon the ranges must cover and
    ``prepare(lattice)`` is applied to the calibration and the "me16"
    lattice alike.  ``reference_arithmetic=True`` gives an "me16ref"
    lattice: the same codes, stepped in the reference variant's FP64
    arithmetic (decode, bgk_step, y-walls, encode), periodic or y-walls."""
    if ranges is None:
        ranges = calibrate_moment_ranges(shape, tau, steps, margin=margin, warmup_steps=warmup_steps,
                                         populations=populations, rho=rho, u=u, prepare=prepare, **lattice_kw)
    lat = Lattice(shape, tau, storage="me16ref" if reference_arithmetic else "me16", **lattice_kw)
    lat.set_moment_ranges(*ranges)
    if prepare is not None:
        prepare(lat)
    return _load_state(lat, populations, rho, u)
