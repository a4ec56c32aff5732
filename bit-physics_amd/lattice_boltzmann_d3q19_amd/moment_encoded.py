"""16-bit moment-encoded storage ("me16", "me16ref"): range calibration and set-up.

The reference's frontier-moment-encoded variant
(packages/lattice-boltzmann-d3q19-frontier-moment-encoded/src/lbm_me.cpp:222-253,
``calibrate_ranges``) quantizes each of the 19 moments over a range taken from
an FP64 run of the whole trajectory: the minimum and maximum of the moment
over every cell of every frame 0 .. max(warmup_steps, steps), widened on both
sides by ``range_margin`` times the span (a span below 1e-12 counts as
1e-12), so that no moment clamps.  :func:`calibrate_moment_ranges` runs that
FP64 trajectory on the GPU (an "f64" lattice of the same configuration,
``Lattice.moment_range`` per frame); :func:`moment_encoded_lattice` returns a
lattice with those ranges and the initial state loaded.

The variant encodes f^n, the state after streaming ("me16ref" does the
same); "me16" encodes what the engine stores, the post-collision populations
of the last step (f^0 before the first), whose non-conserved moments
overshoot f^n's for tau < 1.  Its ranges are therefore taken over the stored
states of the FP64 run (``moment_range(stored=True)``), frames 0 .. T.
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
                            populations=None, rho=None, u=None, prepare=None, reference_arithmetic=False,
                            **lattice_kw):
    """(lo, hi) of the 19 moments over the FP64 trajectory of ``steps`` steps
    (at least ``warmup_steps``) from the given initial state, padded: of the
    stored states for an "me16" lattice, of f^n (the variant's rule) with
    ``reference_arithmetic`` for an "me16ref" one.

    ``lattice_kw`` are the Lattice arguments of the run (boundary, walls,
    force, ...) and ``prepare(lattice)`` sets what they do not (solids, a
    Brinkman field, a force field); the moment-encoded lattice must be
    configured the same way.
    """
    stored = not reference_arithmetic
    lat = Lattice(shape, tau, storage="f64", **lattice_kw)
    if prepare is not None:
        prepare(lat)
    lat = _load_state(lat, populations, rho, u)
    lo, hi = lat.moment_range(stored=stored)
    for _ in range(max(int(warmup_steps), int(steps))):
        lat.step(1)
        a, b = lat.moment_range(stored=stored)
        lo = np.minimum(lo, a)
        hi = np.maximum(hi, b)
    lat.close()
    return pad_ranges(lo, hi, margin)


def moment_encoded_lattice(shape, tau, steps, *, ranges=None, margin=RANGE_MARGIN, warmup_steps=WARMUP_STEPS,
                           populations=None, rho=None, u=None, prepare=None, reference_arithmetic=False,
                           **lattice_kw):
    """An "me16" lattice with calibrated (or given) ranges and the initial
    state loaded; ``steps`` is the horizon the ranges must cover and
    ``prepare(lattice)`` is applied to the calibration and the lattice
    alike.  ``reference_arithmetic=True`` gives an "me16ref" lattice: codes
    of f^n stepped in the reference variant's FP64 arithmetic (decode,
    collide, stream + y-walls, encode), periodic or y-walls, with ranges
    over f^n as the variant takes them."""
    if ranges is None:
        ranges = calibrate_moment_ranges(shape, tau, steps, margin=margin, warmup_steps=warmup_steps,
                                         populations=populations, rho=rho, u=u, prepare=prepare,
                                         reference_arithmetic=reference_arithmetic, **lattice_kw)
    lat = Lattice(shape, tau, storage="me16ref" if reference_arithmetic else "me16", **lattice_kw)
    lat.set_moment_ranges(*ranges)
    if prepare is not None:
        prepare(lat)
    return _load_state(lat, populations, rho, u)
