"""GPU parity of the moment-encoded solver's C++ interface (libd3q19_me.so).

* the C++ test program (tests/cpp/test_lbm_me.cpp): the reference's acceptance
  cases for this interface (packages/lattice-boltzmann-d3q19-frontier-moment-
  encoded/tests/test_lbm_me.cpp:47-164) and the frame schedule / errors;
* reference_step of a perturbed state against the oracle's reference loop
  (bgk_step + apply_bounce_back_y_walls, force (force_x, 0, 0));
* the C++ captures (16-bit codes and FP64) against the Python Lattice making
  the same C-ABI calls (bit for bit: the same kernels on the same inputs), the
  FP64 capture against the oracle trajectory, and the witness against hashlib
  of the final populations;
* the manifest against the capture-v1 rules, the payload checksum, and the
  lbm_me_capture tool (lbm_me_capture_main.cpp:10-31).
"""

from __future__ import annotations

import hashlib
import json
import re
import subprocess

import numpy as np
import pytest

import oracle
from conftest import ROOT, field_scaled_err, print_case

pytestmark = pytest.mark.gpu

TEST_BIN = ROOT / "tests" / "cpp" / "build" / "test_lbm_me"
TOOL = ROOT / "bit-physics_amd" / "bin" / "lbm_me_capture"
SHAPE, TAU, FX, STEPS, EVERY = (64, 32, 3), 0.7, 1e-5, 60, 7
FRAME_STEPS = [0, 7, 14, 21, 28, 35, 42, 49, 56, 60]


@pytest.fixture(scope="module")
def dump(tmp_path_factory):
    print_case("cpp-me-test-program")
    d = tmp_path_factory.mktemp("cpp_me")
    r = subprocess.run([str(TEST_BIN), str(d)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:], r.stderr[-3000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.rstrip().endswith("all passed (0 failed checks)")
    return d


def results(d):
    lines = (d / "results.txt").read_text().split("\n")
    runs = []
    for line in lines[:2]:
        w, m0, mT, pxT, px0, nf = line.split()
        runs.append({"witness": w, "mass0": float(m0), "massT": float(mT), "pxT": float(pxT), "px0": float(px0),
                     "frames": int(nf)})
    rng = np.array([[float(v) for v in line.split()] for line in lines[2:21]])
    return runs, rng[:, 0], rng[:, 1]


def test_cpp_acceptance_cases(dump):
    assert (dump / "results.txt").exists()


def test_reference_step_matches_oracle(dump):
    print_case("cpp-me-reference-step")
    with np.load(dump / "reference_step.npz", allow_pickle=False) as z:
        f_in, f_out, tau, fx = z["f_in"], z["f_out"], float(z["tau"]), float(z["force_x"])
    F = np.zeros((3,) + f_in.shape[1:])
    F[0] = fx
    want = oracle.run_reference(f_in, tau, 1, force_lattice=F, ywalls=True)
    assert field_scaled_err(f_out, want) < 1e-14


def rest(shape):
    return np.broadcast_to(oracle.constants()[1][:, None, None, None], (19,) + shape).astype(np.float64).copy()


def python_run(amd, storage, lo=None, hi=None):
    """The C++ run's C-ABI calls made from Python: frames (rho, u) and the final populations."""
    lat = amd.Lattice(SHAPE, TAU, storage=storage, boundary="ywalls", force=(FX, 0.0, 0.0))
    if storage in ("me16", "me16ref"):
        lat.set_moment_ranges(lo, hi)
    lat.set_populations(rest(SHAPE))
    frames, s = {}, 0
    for n in FRAME_STEPS:
        if n > s:
            lat.step(n - s)
        s = n
        frames[n] = lat.moments()
    f = lat.populations()
    lat.close()
    return frames, f


@pytest.mark.parametrize("mode", ["me16", "f64"])
def test_capture_equals_python_lattice(dump, mode):
    import lattice_boltzmann_d3q19_amd as amd
    from lattice_boltzmann_d3q19_amd.capture import load_capture, validate_manifest
    from lattice_boltzmann_d3q19_amd.moment_encoded import calibrate_moment_ranges

    print_case("cpp-me-capture", mode=mode)
    runs, lo, hi = results(dump)
    run = runs[0 if mode == "me16" else 1]
    manifest = dump / f"{mode}_capture.json"
    m = json.loads(manifest.read_text())
    validate_manifest(m)
    assert m["payload"] == {"format": "npz", "path": f"{mode}_capture.npz",
                            "checksum": "sha256:" + hashlib.sha256((dump / f"{mode}_capture.npz").read_bytes()).hexdigest()}
    assert m["sim"]["variant"] == ("frontier-moment-encoded-16bit" if mode == "me16" else "stack-c-f64")
    assert m["config"]["dims"] == list(SHAPE) and m["run"]["step_count"] == STEPS and m["run"]["capture_interval"] == EVERY
    assert m["config"]["params"]["quantize_bits"] == (16.0 if mode == "me16" else 0.0)
    cap = load_capture(manifest)
    assert [s.step for s in cap.steps()] == FRAME_STEPS
    assert cap.metadata["sim_variant"] == m["sim"]["variant"] and int(cap.metadata["seed"]) == 42
    if mode == "me16":
        plo, phi = calibrate_moment_ranges(SHAPE, TAU, STEPS, populations=rest(SHAPE), boundary="ywalls",
                                           force=(FX, 0.0, 0.0), reference_arithmetic=True)
        assert np.array_equal(plo, lo) and np.array_equal(phi, hi)
        # run_lbm steps the codes in the reference variant's FP64 arithmetic
        frames, f = python_run(amd, "me16ref", lo, hi)
    else:
        frames, f = python_run(amd, "f64")
    for n in FRAME_STEPS:
        st = cap.step(n)
        rho, u = frames[n]
        assert np.array_equal(st.state["rho"], rho.astype(np.float64)), n
        assert np.array_equal(st.state["u"], u.astype(np.float64)), n
        assert st.diagnostics["u_max_lat"] == float(np.abs(st.state["u"]).max())
        assert st.diagnostics["rho_min"] == float(st.state["rho"].min())
    assert run["witness"] == hashlib.sha256(np.ascontiguousarray(f, dtype="<f8").tobytes()).hexdigest()
    assert run["frames"] == len(FRAME_STEPS)
    r0, rT = cap.step(0).state["rho"], cap.step(STEPS).state["rho"]
    # the C++ side sums in cell order, numpy pairwise
    assert abs(run["mass0"] - r0.sum()) <= 1e-12 * r0.size
    assert abs(run["massT"] - rT.sum()) <= 1e-12 * rT.size


def test_f64_capture_is_the_reference_loop(dump):
    from lattice_boltzmann_d3q19_amd.capture import load_capture

    print_case("cpp-me-f64-vs-oracle")
    cap = load_capture(dump / "f64_capture.json")
    F = np.zeros((3,) + SHAPE)
    F[0] = FX
    f, s = rest(SHAPE), 0
    for n in FRAME_STEPS:
        f = oracle.run_reference(f, TAU, n - s, force_lattice=F, ywalls=True)
        s = n
        st = cap.step(n)
        assert field_scaled_err(st.state["rho"], oracle.density_field(f)) < 1e-13, n
        assert field_scaled_err(st.state["u"], oracle.macroscopic_velocity(f, F)) < 1e-10, n


def test_me16_capture_is_the_oracle_quantized_trajectory(dump):
    """run_lbm(quantize=true) steps the codes in the reference variant's FP64
    arithmetic: every captured frame's rho and u are those of the oracle's
    quantized trajectory (decode, bgk_step, y-walls, encode) to the bit, and
    the witness is the SHA-256 of the oracle's final decoded populations."""
    from lattice_boltzmann_d3q19_amd.capture import load_capture

    print_case("cpp-me-me16-vs-oracle")
    runs, lo, hi = results(dump)
    cap = load_capture(dump / "me16_capture.json")
    F = np.zeros((3,) + SHAPE)
    F[0] = FX
    for n in FRAME_STEPS:
        q = oracle.run_me(rest(SHAPE), TAU, n, lo, hi, force_lattice=F, ywalls=True)
        st = cap.step(n)
        assert np.array_equal(st.state["rho"], oracle.density_field(q)), n
        assert np.array_equal(st.state["u"], oracle.macroscopic_velocity(q, F)), n
    assert runs[0]["witness"] == hashlib.sha256(np.ascontiguousarray(q, dtype="<f8").tobytes()).hexdigest()


def test_capture_tool(tmp_path):
    print_case("cpp-me-capture-tool")
    r = subprocess.run([str(TOOL), str(tmp_path / "q.json"), "--steps", "20"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    mt = re.fullmatch(r"frames=21 witness=([0-9a-f]{64}) mass0=(\S+) massT=(\S+)\n", r.stdout)
    assert mt, r.stdout
    assert abs(float(mt.group(3)) - float(mt.group(2))) <= 1e-3 * 64 * 32 * 3
    assert (tmp_path / "q.npz").exists()
    r = subprocess.run([str(TOOL), str(tmp_path / "f.json"), "--f64", "--steps", "5"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("frames=6 "), r.stdout + r.stderr
    assert json.loads((tmp_path / "f.json").read_text())["sim"]["variant"] == "stack-c-f64"
    assert subprocess.run([str(TOOL)], capture_output=True, timeout=60).returncode == 2
    assert subprocess.run([str(TOOL), str(tmp_path / "x.json"), "--steps", "-3"], capture_output=True,
                          timeout=60).returncode == 2
