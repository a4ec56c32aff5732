"""D3Q19 constants and canonical run sizes (reference/constants.py:39-100)."""

from __future__ import annotations

import numpy as np

VELOCITIES = (
    (0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
    (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1),
    (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1),
)
WEIGHTS = (1.0 / 3.0, *([1.0 / 18.0] * 6), *([1.0 / 36.0] * 12))
CS2 = 1.0 / 3.0
C = np.asarray(VELOCITIES, dtype=np.int64)
W = np.asarray(WEIGHTS, dtype=np.float64)
OPP = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17], dtype=np.int64)

CANONICAL_DESCRIPTOR_POISEUILLE = "poiseuille-64x32-seed42-step1000"
CANONICAL_DESCRIPTOR_COUETTE = "couette-32x16-seed42-step500"
CANONICAL_SEED = 42
CANONICAL_NZ = 3
CANONICAL_POISEUILLE_NX = 64
CANONICAL_POISEUILLE_NY = 32
CANONICAL_POISEUILLE_STEPS = 1000
CANONICAL_COUETTE_NX = 32
CANONICAL_COUETTE_NY = 16
CANONICAL_COUETTE_STEPS = 500
