"""Build the C++ test program of libd3q19_me.so: tests/cpp/build/test_lbm_me.

Needs the package built first (bit-physics_amd/build.py makes libd3q19.so and
libd3q19_me.so); links against them with a relative run path, so the tree can
move as a whole.
"""

from __future__ import annotations

import importlib.util
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
OUT = HERE / "build" / "test_lbm_me"


def build(verbose: bool = False) -> Path:
    spec = importlib.util.spec_from_file_location("d3q19_hip_build", ROOT / "bit-physics_amd" / "build.py")
    pkg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(pkg)
    OUT.parent.mkdir(exist_ok=True)
    rel = "$ORIGIN/../../../bit-physics_amd/lib"
    cmd = [pkg.cxx(), "-std=c++17", "-O2", "-Wall", "-Wextra", f"-I{pkg.INCLUDE}", f"-I{pkg.CPP_DIR}",
           str(HERE / "test_lbm_me.cpp"), "-o", str(OUT), f"-L{pkg.LIB_DIR}", "-ld3q19_me", *pkg.cpp_link_flags(rel)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
