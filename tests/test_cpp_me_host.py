"""Host side of the moment-encoded C++ interface (libd3q19_me.so), no GPU.

The C++ test program's --host-only mode writes SHA-256 digests, an .npz
archive and a JSON document; here hashlib, numpy.load (allow_pickle=False)
and json read them back.  The header is checked to declare the reference's
interface (packages/lattice-boltzmann-d3q19-frontier-moment-encoded/include/
bit_physics/lbm_d3q19_me/lbm_me.hpp:30-117).
"""

from __future__ import annotations

import hashlib
import json
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

TEST_BIN = ROOT / "tests" / "cpp" / "build" / "test_lbm_me"
HEADER = ROOT / "include" / "bit_physics" / "lbm_d3q19_me" / "lbm_me.hpp"


@pytest.fixture(scope="module")
def host_dir(tmp_path_factory):
    if not TEST_BIN.exists():
        pytest.fail(f"{TEST_BIN} is not built (run __graft_entry__.build())")
    d = tmp_path_factory.mktemp("cpp_host")
    r = subprocess.run([str(TEST_BIN), "--host-only", str(d)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host-only: ok" in r.stdout
    return d


def test_sha256_matches_hashlib(host_dir):
    ramp = bytes((k * 7 + 3) & 0xFF for k in range(200))
    lines = (host_dir / "sha256.txt").read_text().split("\n")
    got = dict(line.split() for line in lines if line)
    assert len(got) == 201
    for n in range(201):
        assert got[str(n)] == hashlib.sha256(ramp[:n]).hexdigest(), n


def test_npz_reads_back(host_dir):
    with np.load(host_dir / "host.npz", allow_pickle=False) as z:
        assert z.files == ["steps/3/state/a", "line", "scalar", "count", "name", "empty"]
        a = 0.5 * np.arange(6) - 1.25
        assert np.array_equal(z["steps/3/state/a"], a.reshape(2, 3)) and z["steps/3/state/a"].dtype == np.float64
        assert np.array_equal(z["line"], a) and z["line"].shape == (6,)
        assert z["scalar"].shape == () and float(z["scalar"]) == 3.25
        assert z["count"].dtype == np.int64 and int(z["count"]) == -42
        assert z["name"].item() == "lattice-boltzmann-d3q19"
        assert z["empty"].shape == (0,)


def test_json_reads_back(host_dir):
    d = json.loads((host_dir / "host.json").read_text())
    assert d == {"a": [64, True, 'quote " back \\ tab \t'], "b": {"x": 0.7, "y": 16.0, "z": 1e-5}, "e": {}}
    assert list(d) == sorted(d)


def test_header_declares_the_reference_interface():
    text = HEADER.read_text()
    assert "namespace bit_physics::lbm_d3q19_me" in text
    for sig in (
        r"MomentBasis build_moment_basis\(\);",
        r"QuantRanges calibrate_ranges\(const LbmConfig& cfg, const MomentBasis& basis\);",
        r"void reference_step\(std::vector<double>& f, const LbmConfig& cfg\);",
        r"std::vector<double> initial_rest_state\(const LbmConfig& cfg\);",
        r"LbmResult run_lbm\(const LbmConfig& cfg, const std::filesystem::path\* capture_manifest\);",
        r"std::array<double, kQ> bound\(\) const;",
    ):
        assert re.search(sig, text), sig
    # LbmConfig defaults of the reference (lbm_me.hpp:47-59)
    for field in ("nx = 64, ny = 32, nz = 3", "steps = 1000", "capture_interval = 1", "tau = 0.7",
                  "force_x = 1e-5", "seed = 42", "quantize = true", "range_margin = 0.25", "warmup_steps = 64"):
        assert field in text, field
    for field in ("determinism_witness_sha256", "total_mass_initial", "total_mass_final", "total_momentum_initial",
                  "total_momentum_final", "ranges_used", "inverse_residual", "m_inv"):
        assert field in text, field


def test_library_exports_the_interface():
    lib = ROOT / "bit-physics_amd" / "lib" / "libd3q19_me.so"
    out = subprocess.run(["nm", "-DC", "--defined-only", str(lib)], capture_output=True, text=True, check=True).stdout
    for name in ("bit_physics::lbm_d3q19_me::build_moment_basis()",
                 "bit_physics::lbm_d3q19_me::calibrate_ranges(", "bit_physics::lbm_d3q19_me::reference_step(",
                 "bit_physics::lbm_d3q19_me::initial_rest_state(", "bit_physics::lbm_d3q19_me::run_lbm(",
                 "bit_physics::lbm_d3q19_me::QuantRanges::bound() const"):
        assert name in out, name
