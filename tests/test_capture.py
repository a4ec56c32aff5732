"""Capture files, manifest rules and run diagnostics, on the CPU.

Mirrors the reference testkit's capture and determinism tests
(tools/testkit/capture/tests, packages/lattice-boltzmann-d3q19/tests/
test_determinism.py:70-130): round trips, the manifest rules, the
bit-exact diff and its drift witness.  The diagnostics are checked against
loop-by-loop evaluations of their definitions.
"""

from __future__ import annotations

import copy
import json

import numpy as np
import pytest

from lattice_boltzmann_d3q19_amd.capture import (
    CaptureManifest,
    ManifestError,
    StepState,
    compare_captures,
    diff_captures,
    load_capture,
    run_twice_and_diff,
    validate_manifest,
    write_capture,
)
from lattice_boltzmann_d3q19_amd.sim import health_check, vector_field_diagnostics


def manifest(path="stub.h5", seed=42, dims=(4,)):
    return CaptureManifest(
        schema_version="1.0.0",
        sim={"name": "lattice-boltzmann-d3q19", "category": "lattice", "variant": "bgk-d3q19-qian-1992"},
        stack={"name": "stub", "version": "0.0.1", "build_id": "stub"},
        config={"tier": "test", "dims": list(dims), "dtype": "f64", "seed": seed, "params": {}},
        run={"step_count": 2, "capture_interval": 1, "wall_clock_seconds": 0.0, "start_utc": "2026-01-01T00:00:00Z"},
        payload={"format": "hdf5", "path": path, "checksum": "sha256:" + "0" * 64},
        determinism={"claimed": "bit-exact-same-hw", "atomic_ops": False, "subgroup_ops": False},
    )


def frames(shift=0.0):
    rng = np.random.default_rng(3)
    out = []
    for n in (0, 1, 2):
        rho = 1.0 + 1e-3 * rng.standard_normal((4, 3, 2))
        u = 1e-2 * rng.standard_normal((3, 4, 3, 2))
        u[0, 0, 0, 0] += shift
        out.append(StepState(step=n, state={"rho": rho, "u": u}, diagnostics={"check_health": 0.0, "x": n * 0.5}))
    return out


def test_manifest_rules():
    good = manifest().to_dict()
    validate_manifest(good)
    for mutate in (
        lambda m: m.pop("sim"),
        lambda m: m["sim"].update(extra=1),
        lambda m: m.update(extra=1),
        lambda m: m["config"].update(dtype="f16"),
        lambda m: m["config"].update(dims=[]),
        lambda m: m["config"].update(dims=[0]),
        lambda m: m["config"].update(seed=True),
        lambda m: m["run"].update(capture_interval=0),
        lambda m: m["run"].update(wall_clock_seconds=-1.0),
        lambda m: m["payload"].update(format="zarr"),
        lambda m: m["payload"].update(checksum="sha256:xyz"),
        lambda m: m["determinism"].update(claimed="mostly"),
        lambda m: m["determinism"].update(atomic_ops=0),
        lambda m: m.update(schema_version="1.0"),
        lambda m: m["stack"].update(name=""),
    ):
        bad = copy.deepcopy(good)
        mutate(bad)
        with pytest.raises(ManifestError):
            validate_manifest(bad)


def test_npz_round_trip(tmp_path):
    fr = frames()
    path = write_capture(fr, manifest(), tmp_path, payload_format="npz")
    m = json.loads(path.read_text())
    assert m["payload"]["format"] == "npz" and m["payload"]["path"] == "stub.npz"
    assert m["payload"]["checksum"].startswith("sha256:") and m["payload"]["checksum"] != "sha256:" + "0" * 64
    cap = load_capture(path)
    assert [s.step for s in cap.steps()] == [0, 1, 2]
    for a in fr:
        b = cap.step(a.step)
        for k in ("rho", "u"):
            assert np.array_equal(a.state[k], b.state[k]) and b.state[k].dtype == np.float64
        assert b.diagnostics == a.diagnostics
    assert cap.metadata["sim_name"] == "lattice-boltzmann-d3q19" and int(cap.metadata["seed"]) == 42
    assert np.array_equal(cap.field(2, "u"), fr[2].state["u"])
    with pytest.raises(KeyError):
        cap.step(7)


def test_diff_and_drift_witness(tmp_path):
    a = write_capture(frames(), manifest(), tmp_path / "a", payload_format="npz")
    b = write_capture(frames(), manifest(), tmp_path / "b", payload_format="npz")
    c = write_capture(frames(shift=1e-9), manifest(), tmp_path / "c", payload_format="npz")
    assert diff_captures(a, b).bit_exact
    d = diff_captures(a, c)
    assert not d.bit_exact and d.max_abs_err == pytest.approx(1e-9, rel=1e-3)
    assert all(f.endswith(":u") for f in d.mismatched_fields) and len(d.mismatched_fields) == 3
    # the lbm cross-stack tolerance (relative 1e-5 of max |u|) accepts this drift
    assert compare_captures(a, c).bit_exact
    calls = {"n": 0}

    def drifting(seed, out_dir):
        calls["n"] += 1
        return write_capture(frames(shift=calls["n"]), manifest(seed=seed), out_dir, payload_format="npz")

    v = run_twice_and_diff(drifting, seed=42, tmp_dir=tmp_path / "det")
    assert not v.content_equivalent and "max_abs_err" in v.detail
    v = run_twice_and_diff(lambda s, o: write_capture(frames(), manifest(seed=s), o, payload_format="npz"),
                           seed=42, tmp_dir=tmp_path / "det2")
    assert v.content_equivalent and v.detail == "captures match exactly"


def test_step_set_mismatch(tmp_path):
    a = write_capture(frames(), manifest(), tmp_path / "a", payload_format="npz")
    b = write_capture(frames()[:2], manifest(), tmp_path / "b", payload_format="npz")
    d = diff_captures(a, b)
    assert not d.bit_exact and d.mismatched_fields == ["step:2:only-left"]


def test_health_check_definition():
    rng = np.random.default_rng(5)
    rho = 1 + 0.01 * rng.standard_normal((5, 4, 3))
    u = 0.02 * rng.standard_normal((3, 5, 4, 3))
    h = health_check(rho, u)
    umax = max(np.sqrt(u[0][i] ** 2 + u[1][i] ** 2 + u[2][i] ** 2) for i in np.ndindex(5, 4, 3))
    assert h == {"check_health": 0.0, "rho_min": rho.min(), "rho_max": rho.max(), "u_max_lat": pytest.approx(umax, rel=1e-15)}
    u[1, 2, 1, 0] = np.nan
    assert health_check(rho, u)["check_health"] == 1.0
    rho[0, 0, 0] = np.inf
    assert health_check(rho, np.zeros_like(u))["check_health"] == 1.0


def test_vector_field_definition():
    rng = np.random.default_rng(6)
    nx, ny, nz, dx = 6, 5, 3, 1.0 / 6
    u = rng.standard_normal((3, nx, ny, nz))
    got = vector_field_diagnostics(u, dx)
    best = 0.0
    for i in range(nx):
        for j in range(ny):
            div = (u[0, (i + 1) % nx, j, 0] - u[0, (i - 1) % nx, j, 0]) / (2 * dx) + \
                  (u[1, i, (j + 1) % ny, 0] - u[1, i, (j - 1) % ny, 0]) / (2 * dx)
            best = max(best, abs(div))
    assert got["check_divergence_free"] == pytest.approx(best, rel=1e-13)
    assert got["check_circulation"] == pytest.approx(u[0, :, :, 0].sum() * dx * dx, rel=1e-13)
    # a divergence-free periodic field: u = (sin(2 pi y / ny), 0, 0)
    y = np.arange(ny)
    v = np.zeros((3, nx, ny, nz))
    v[0] = np.sin(2 * np.pi * y / ny)[None, :, None]
    assert vector_field_diagnostics(v, dx)["check_divergence_free"] == 0.0


def test_npz_payload_bytes_are_reproducible(tmp_path):
    """Equal frames give byte-identical payloads (fixed member timestamps)."""
    import time

    a = write_capture(frames(), manifest(), tmp_path / "a", payload_format="npz")
    time.sleep(2.1)  # a zip timestamp has 2 s resolution
    b = write_capture(frames(), manifest(), tmp_path / "b", payload_format="npz")
    ma, mb = json.loads(a.read_text()), json.loads(b.read_text())
    assert ma["payload"]["checksum"] == mb["payload"]["checksum"]
    assert (tmp_path / "a" / "stub.npz").read_bytes() == (tmp_path / "b" / "stub.npz").read_bytes()
    c = write_capture(frames(shift=1e-9), manifest(), tmp_path / "c", payload_format="npz")
    assert json.loads(c.read_text())["payload"]["checksum"] != ma["payload"]["checksum"]
