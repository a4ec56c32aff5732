"""GPU runs of the canonical runners, checked like the reference checks its own.

* tier-1/2 diagnostics of the diagnostic-tier run
  (packages/lattice-boltzmann-d3q19/tests/test_diagnostics.py:28-80);
* run-twice content equivalence (tests/test_determinism.py:53-67);
* the published known answers through the runners: Poiseuille step 1000
  max|u| = 8.6508320e-03, max rho = 1.0000000; Couette step 500
  max|u| = 5.0000000e-02 (SPEC/equivalence.md:240-243);
* every frame of the GPU capture against a capture of the same run made by
  the FP64 oracle: field-scaled 1e-10 with FP64 storage (the engine stores
  f - w and the oracle f, so roundings differ; on the 16 x 8 x 3 Poiseuille
  run the gap is 3.8e-15 absolute on a u field of max 4e-4, i.e. ~1e-11),
  and within the lbm cross-stack tolerance (relative 1e-5) with FP32 storage.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import print_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sim():
    from lattice_boltzmann_d3q19_amd import sim as s

    return s


def oracle_capture(sim, out_dir, shape, n_steps, interval, force_x=0.0, wall=0.0):
    """The same frames as a runner, from the CPU oracle."""
    from lattice_boltzmann_d3q19_amd.capture import StepState, write_capture

    nx = shape[0]
    f = oracle.feq_field(np.ones(shape), np.zeros((3,) + shape))
    F = None
    if force_x:
        F = np.zeros((3,) + shape)
        F[0] = force_x
    frames = []

    def frame(n, f):
        rho = oracle.density_field(f)
        u = oracle.macroscopic_velocity(f, F)
        frames.append(StepState(step=n, state={"rho": rho, "u": u},
                                diagnostics={**sim.health_check(rho, u),
                                             **sim.vector_field_diagnostics(u, 1.0 / nx)}))

    frame(0, f)
    done = 0
    while done < n_steps:
        nxt = min(n_steps, (done // interval + 1) * interval)
        f = oracle.run_reference(f, sim.CANONICAL_TAU, nxt - done, force_lattice=F, ywalls=True,
                                 wall_velocity_top=(wall, 0.0, 0.0))
        done = nxt
        frame(done, f)
    m = sim._manifest("oracle-run", "test", shape, 42, {}, n_steps, interval, "f64")
    return write_capture(frames, m, out_dir, payload_format="npz")


def test_diagnostic_run_health_and_vector_field(sim, tmp_path):
    from lattice_boltzmann_d3q19_amd.capture import load_capture

    print_case("sim-diagnostic")
    cap = load_capture(sim.sim_runner_diagnostic(42, tmp_path, payload_format="npz"))
    steps = [s.step for s in cap.steps()]
    assert steps == [0, 10, 20, 30, 40, 50]
    for st in cap.steps():
        for name, arr in st.state.items():
            assert np.all(np.isfinite(arr)), (name, st.step)
        for name, v in st.diagnostics.items():
            assert np.isfinite(v), (name, st.step)
        assert st.diagnostics["check_health"] == 0.0
    last = cap.step(50).diagnostics
    assert last["check_divergence_free"] < 1.0
    assert cap.manifest.config["tier"] == "diagnostic" and cap.manifest.run["capture_interval"] == 10


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_run_twice_content_equivalent(sim, tmp_path, storage):
    from lattice_boltzmann_d3q19_amd.capture import run_twice_and_diff

    print_case("sim-run-twice", storage=storage)
    v = run_twice_and_diff(lambda s, o: sim.sim_runner_diagnostic(s, o, storage=storage, payload_format="npz"),
                           seed=42, tmp_dir=tmp_path)
    assert v.content_equivalent, v.detail
    assert v.detail == "captures match exactly"


def test_canonical_poiseuille_known_answer(sim, tmp_path):
    from lattice_boltzmann_d3q19_amd.capture import load_capture

    print_case("sim-poiseuille-canonical")
    cap = load_capture(sim.sim_runner_seeded(42, tmp_path, payload_format="npz"))
    assert len(list(cap._step_numbers())) == 1001
    last = cap.step(1000)
    umax = float(np.max(np.sqrt((last.state["u"] ** 2).sum(0))))
    assert f"{umax:.7e}" == "8.6508320e-03"
    assert f"{float(last.state['rho'].max()):.7f}" == "1.0000000"
    assert last.diagnostics["u_max_lat"] == umax


def test_canonical_couette_known_answer(sim, tmp_path):
    from lattice_boltzmann_d3q19_amd.capture import load_capture

    print_case("sim-couette-canonical")
    cap = load_capture(sim.sim_runner_seeded_couette(42, tmp_path, payload_format="npz"))
    last = cap.step(500)
    umax = float(np.max(np.sqrt((last.state["u"] ** 2).sum(0))))
    assert f"{umax:.7e}" == "5.0000000e-02"


@pytest.mark.parametrize("storage,rtol", [("f64", 1e-10), ("f32", 1e-5)])
def test_diagnostic_capture_matches_oracle_capture(sim, tmp_path, storage, rtol):
    from lattice_boltzmann_d3q19_amd.capture import compare_captures

    print_case("sim-vs-oracle", storage=storage)
    ours = sim.sim_runner_diagnostic(42, tmp_path / "gpu", storage=storage, payload_format="npz")
    ref = oracle_capture(sim, tmp_path / "cpu", (16, 8, 3), 50, 10, force_x=1e-5)
    d = compare_captures(ours, ref, rtol=rtol)
    assert d.bit_exact, (d.max_abs_err, d.max_rel_err, d.mismatched_fields[:3])


def test_couette_capture_matches_oracle_capture(sim, tmp_path):
    from lattice_boltzmann_d3q19_amd.capture import compare_captures, write_capture

    print_case("sim-couette-vs-oracle")
    lat = sim.couette_lattice(32, 16, 3, storage="f64")
    m = sim._manifest("gpu-couette", "test", (32, 16, 3), 42, {}, 100, 25, "f64")
    ours = write_capture(sim.evolve(lat, 100, 25, 1.0 / 32), m, tmp_path / "gpu", payload_format="npz")
    ref = oracle_capture(sim, tmp_path / "cpu", (32, 16, 3), 100, 25, wall=sim.CANONICAL_COUETTE_WALL_VELOCITY)
    d = compare_captures(ours, ref, rtol=1e-10)
    assert d.bit_exact, (d.max_abs_err, d.max_rel_err, d.mismatched_fields[:3])


def _locked_manifest():
    """The diagnostic-tier manifest this engine emits, minus its volatile
    fields (packages/lattice-boltzmann-d3q19/tests/test_manifest_equality.py:
    the run's wall clock, its start time and the payload checksum)."""
    return {
        "schema_version": "1.0.0",
        "sim": {"name": "lattice-boltzmann-d3q19", "category": "lattice", "variant": "bgk-d3q19-qian-1992"},
        "stack": {"name": "mi355x-hip", "version": "0.1.0", "build_id": "libd3q19-gfx950"},
        "config": {"tier": "diagnostic", "dims": [16, 8, 3], "dtype": "f64", "seed": 42,
                   "params": {"descriptor": "poiseuille-16x8-seed42-step50", "tau": 0.7, "force_x_lattice": 1.0e-5,
                              "nz_convention": "depth-3-z-periodic-slab",
                              "boundary": "bounce-back-y-walls-periodic-xz", "storage": "f64"}},
        "run": {"step_count": 50, "capture_interval": 10},
        "payload": {"format": "npz", "path": "poiseuille-16x8-seed42-step50.npz"},
        "determinism": {"claimed": "bit-exact-same-hw", "atomic_ops": False, "subgroup_ops": False},
    }


def test_diagnostic_manifest_locked_and_stable(sim, tmp_path):
    import json

    print_case("sim-manifest-lock")
    got = []
    for k in range(2):
        m = json.loads(sim.sim_runner_diagnostic(42, tmp_path / str(k), payload_format="npz").read_text())
        assert m["run"]["wall_clock_seconds"] >= 0 and m["run"]["start_utc"].endswith("Z")
        assert m["payload"]["checksum"].startswith("sha256:") and len(m["payload"]["checksum"]) == 71
        got.append(m)
        volatile = {"wall_clock_seconds", "start_utc"}
        locked = {k2: (v if k2 != "run" else {a: b for a, b in v.items() if a not in volatile}) for k2, v in m.items()}
        locked["payload"] = {a: b for a, b in m["payload"].items() if a != "checksum"}
        assert locked == _locked_manifest()
    # run to run the deterministic subset, checksum included, is identical
    assert got[0]["payload"]["checksum"] == got[1]["payload"]["checksum"]
