"""Host-only test of the launch-shape arithmetic in
bit-physics_amd/csrc/d3q19_launch.hpp (tests/cpp/test_launch.cpp, built with
the host C++ compiler): the LDS reservation that caps the fused step's
resident blocks per CU leaves room for exactly the intended number, the
plane chunks a large step is launched in cover the lattice once, the ring gap
of the moment-coded step, and the adjoint's memory plan picks the checkpoint
interval that needs the least memory."""

from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_reserve_lds_leaves_room_for_exactly_k_blocks(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = tmp_path / "test_launch"
    subprocess.run([cxx, "-std=c++17", "-O2", "-Wall", "-Wextra", f"-I{ROOT / 'bit-physics_amd' / 'csrc'}",
                    str(ROOT / "tests" / "cpp" / "test_launch.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout
