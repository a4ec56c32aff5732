"""bench.py's command line, without a GPU: one GPU only, and the flags exist."""

from __future__ import annotations

import os
import subprocess
import sys

from conftest import ROOT


def run(*args, env=None):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, env=env)


def test_more_than_one_gpu_exits_with_a_message():
    r = run("--gpus", "2")
    assert r.returncode == 2 and r.stdout == ""
    assert "one GPU" in r.stderr
    env = dict(os.environ, WORLD_SIZE="4")
    r = run(env=env)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_help_lists_the_contract_flags():
    r = run("--help")
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup", "--study"):
        assert flag in r.stdout
    for study in ("turbulence", "inverse", "channel", "mms", "cavity", "transient"):
        assert study in r.stdout
