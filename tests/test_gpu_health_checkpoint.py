"""GPU health reduction and checkpoint / restore.

* health (the runners' tier-1 check, NP/sim.py:135-145, in one pass on the
  GPU): equals the host check on the read-back fields; a NaN or Inf poked
  into populations is counted (the failure-detection test the reference's
  design calls for), in every storage;
* checkpoint / restore of the stored state: a run checkpointed at step 40
  (and at step 0, before its first in-place collision) and restored into a
  fresh lattice continues bit for bit like the uninterrupted run, in every
  storage, with a periodic force field, moving reference y-walls, and
  moving half-way walls with solids.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import print_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import lattice_boltzmann_d3q19_amd as m

    return m


def seeded(shape, seed):
    rng = np.random.default_rng(seed)
    f = oracle.feq_field(1 + 0.01 * rng.standard_normal(shape), 0.02 * rng.standard_normal((3,) + shape))
    return f * (1 + 0.01 * rng.standard_normal(f.shape)), 1e-4 * rng.standard_normal((3,) + shape)


@pytest.mark.parametrize("storage", ["f64", "f32", "f16", "bf16", "bf16sr"])
@pytest.mark.parametrize("shape", [(9, 7, 5), (4, 3, 64)], ids=lambda s: "x".join(map(str, s)))
def test_health_matches_host_check_and_counts_nonfinite(amd, storage, shape):
    from lattice_boltzmann_d3q19_amd.sim import health_check

    print_case("health", storage=storage, shape=shape)
    f, _ = seeded(shape, 3)
    lat = amd.Lattice(shape, 0.7, storage=storage, boundary="ywalls", force=(1e-5, 0, 0))
    lat.set_populations(f).step(7)
    rho, u = lat.moments()
    want = health_check(rho, u)
    got = lat.health()
    assert got["check_health"] == 0.0 and got["nonfinite_cells"] == 0
    for k in ("rho_min", "rho_max", "u_max_lat"):
        assert got[k] == pytest.approx(want[k], rel=1e-6 if storage != "f64" else 1e-12)
    bad = f.copy()
    bad[3, 1, 2, 3] = np.nan
    bad[7, 2, 0, 1] = np.inf
    lat2 = amd.Lattice(shape, 0.7, storage=storage).set_populations(bad)
    h = lat2.health()
    assert h["check_health"] == 1.0 and h["nonfinite_cells"] == 2
    # a pull spreads the poisoned populations to their neighbours
    assert lat2.step(1).health()["nonfinite_cells"] >= 2


CASES = {
    "periodic-field": dict(force="field"),
    "ywalls": dict(boundary="ywalls", wall_velocity_top=(0.03, 0, 0), force=(1e-5, 0, 0)),
    "halfway-solids": dict(boundary="halfway", wall_faces=("-y", "+y"), face_velocity={"+y": (0.02, 0, 0)}),
}


@pytest.mark.parametrize("storage", ["f64", "f32", "f16", "bf16", "bf16sr", "me16"])
@pytest.mark.parametrize("case", list(CASES))
def test_checkpoint_restore_continues_bitwise(amd, tmp_path, storage, case):
    print_case("checkpoint", storage=storage, case=case)
    shape = (10, 8, 6)
    f, F = seeded(shape, 5)
    kw = dict(CASES[case])
    field = kw.pop("force", None) == "field"
    if field:
        kw["force"] = F
    solid = np.zeros(shape, bool)
    solid[4:6, 3:5, 2:4] = case == "halfway-solids"

    ranges = None
    if storage == "me16":
        m = oracle.me_moments(f).reshape(19, -1)
        span = m.max(1) - m.min(1) + 1e-3
        ranges = (m.min(1) - span, m.max(1) + span)

    def make():
        lat = amd.Lattice(shape, 0.65, storage=storage, **kw)
        if case == "halfway-solids":
            lat.set_solid(solid)
        if ranges is not None:
            lat.set_moment_ranges(*ranges)
        return lat

    full = make().set_populations(f)
    full.step(80)
    for at in (0, 40):
        first = make().set_populations(f)
        first.step(at)
        first.checkpoint(tmp_path / "ck.npz")
        second = make().restore(tmp_path / "ck.npz")
        assert second.step_count == at
        second.step(80 - at)
        assert second.step_count == 80
        assert np.array_equal(second.populations(np.float64), full.populations(np.float64)), at


def test_restore_rejects_other_lattice(amd, tmp_path):
    print_case("checkpoint-mismatch")
    f, _ = seeded((6, 5, 4), 1)
    lat = amd.Lattice((6, 5, 4), 0.7).set_populations(f)
    lat.checkpoint(tmp_path / "ck.npz")
    with pytest.raises(ValueError):
        amd.Lattice((6, 5, 5), 0.7).restore(tmp_path / "ck.npz")
    with pytest.raises(ValueError):
        amd.Lattice((6, 5, 4), 0.7, storage="f16").restore(tmp_path / "ck.npz")
    # moment codes mean the saved state only under the ranges they were made with
    m = oracle.me_moments(f).reshape(19, -1)
    lo, hi = m.min(1) - 0.1, m.max(1) + 0.1
    me = amd.Lattice((6, 5, 4), 0.7, storage="me16").set_moment_ranges(lo, hi).set_populations(f)
    me.checkpoint(tmp_path / "me.npz")
    same = amd.Lattice((6, 5, 4), 0.7, storage="me16").set_moment_ranges(lo, hi).restore(tmp_path / "me.npz")
    assert np.array_equal(same.populations(), me.populations())
    with pytest.raises(ValueError, match="moment ranges"):
        amd.Lattice((6, 5, 4), 0.7, storage="me16").set_moment_ranges(lo, hi + 0.1).restore(tmp_path / "me.npz")


def test_lattice_on_callers_stream(amd):
    """A lattice on a stream of the caller's own runs as one on torch's current
    stream, bit for bit, although its device inputs (and their converted
    copies) are freed as soon as each call returns and torch reuses the memory."""
    import torch

    print_case("callers-stream")
    shape = (8, 6, 70)
    f, F = seeded(shape, 9)

    def inputs():
        return (torch.tensor(f, device="cuda", dtype=torch.float32),
                torch.tensor(F, device="cuda", dtype=torch.float32))

    f32, F32 = inputs()
    ref = amd.Lattice(shape, 0.62, storage="f16", force=F32).set_populations(f32.to(torch.float64)).step(12)
    s = torch.cuda.Stream()
    f32, F32 = inputs()
    f64 = f32.to(torch.float64)
    s.wait_stream(torch.cuda.current_stream())
    lat = amd.Lattice(shape, 0.62, storage="f16", force=F32, stream=s.cuda_stream)
    del F32
    lat.set_populations(f64)
    del f32, f64
    # freed blocks handed out again on torch's stream and overwritten
    junk = [torch.full((19,) + shape, float("nan"), device="cuda", dtype=torch.float64) for _ in range(2)]
    lat.step(12)
    del junk
    assert np.array_equal(lat.populations(), ref.populations())
    ref.close()
    lat.close()
