"""The differentiable configurations at their full sizes, through the studies
that bench.py runs (``--study inverse``, ``--study channel``).

* Inverse problem, 128^3 periodic, tau 0.6, 2000 steps, checkpoints every 50:
  dL/d(u0, F) along 16 seeded unit directions against central differences of
  the FP64 forward (step 1e-3).  The FP32 adjoint lies within 1e-4 of the
  largest directional derivative (measured 3.1e-5), the FP64 adjoint within
  1e-8 (1.1e-10), and the backward pass costs at most 3x the forward.
* Channel, 256 x 128 x 128, walls at y = 0 and 127, Brinkman obstacles (5%
  volume), 4000 steps, checkpoint interval autotuned to 64 GB: dJ/ds of the
  mean outlet velocity with FP16 and FP32 checkpoints agrees to 1e-4
  (2.7e-6), and matches FP64 central differences along 4 unit directions to
  1e-4 (4.4e-5).
"""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def bench():
    sys.path.insert(0, str(ROOT))
    import bench

    return bench


def test_inverse_problem_gradients_match_central_differences(bench):
    import torch

    print_case("full-size-inverse", shape=(128, 128, 128), steps=2000, K=50)
    r = bench.study_inverse(torch)
    print(f"inverse: FD error FP32 {r['fd_max_err_f32_adjoint']:.3e}, FP64 {r['fd_max_err_f64_adjoint']:.3e}, "
          f"backward/forward {r['timing']['adjoint_forward_ratio']}", flush=True)
    assert r["fd_max_err_f32_adjoint"] < 1e-4
    assert r["fd_max_err_f64_adjoint"] < 1e-8
    assert r["timing"]["adjoint_forward_ratio"] <= 3.0


def test_channel_brinkman_gradient_matches_central_differences(bench):
    import torch

    import lattice_boltzmann_d3q19_amd as amd

    print_case("full-size-channel", shape=(256, 128, 128), steps=4000)
    r = bench.study_channel(amd, torch)
    print(f"channel: f16 vs f32 checkpoints {r['grad_rel_err_f16_vs_f32_checkpoints']:.3e}, "
          f"FD {r['fd_max_rel_err']:.3e}, K {r['checkpoints_f16']['K']} / {r['checkpoints_f32']['K']}", flush=True)
    assert r["grad_rel_err_f16_vs_f32_checkpoints"] < 1e-4
    assert r["fd_max_rel_err"] < 1e-4
    assert r["checkpoints_f16"]["memory_GB"] <= 64.0 and r["checkpoints_f32"]["memory_GB"] <= 64.0
