"""Build libd3q19.so (HIP, gfx950) in-tree with hipcc.

    python bit-physics_amd/build.py

Output: bit-physics_amd/lib/libd3q19.so.  Every call recompiles.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB_DIR = HERE / "lib"
LIB = LIB_DIR / "libd3q19.so"
SOURCES = ["d3q19_capi.hip", "d3q19_adjoint.hip", "d3q19_calibration.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # keep the compiler from forming packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm required to build libd3q19.so)")


NODE_DIR = HERE / "node"
NODE_ADDON = NODE_DIR / "build" / "d3q19.node"


def build_node_addon(verbose: bool = False) -> Path | None:
    """Node-API addon over the C ABI (g++ against the system node headers)."""
    inc = next((p for p in (Path("/usr/include/node"), Path("/usr/local/include/node"))
                if (p / "node_api.h").exists()), None)
    cxx = shutil.which("g++")
    if inc is None or cxx is None:
        if verbose:
            print("node headers or g++ not found: skipping the Node-API addon", file=sys.stderr)
        return None
    NODE_ADDON.parent.mkdir(exist_ok=True)
    cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-shared", "-fPIC", f"-I{inc}",
           "-DNODE_GYP_MODULE_NAME=d3q19", str(NODE_DIR / "d3q19_addon.cc"), "-o", str(NODE_ADDON),
           f"-L{LIB_DIR}", "-ld3q19", "-Wl,-rpath,$ORIGIN/../../lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return NODE_ADDON


CPP_DIR = HERE / "cpp"
INCLUDE = HERE.parent / "include"
ME_LIB = LIB_DIR / "libd3q19_me.so"
BIN_DIR = HERE / "bin"
ROCM_LIB = Path("/opt/rocm/lib")


def cxx() -> str:
    exe = shutil.which("g++")
    if exe is None:
        raise RuntimeError("g++ not found (needed for the C++ interface libd3q19_me.so)")
    return exe


def cpp_link_flags(rpath: str) -> list[str]:
    """Link flags of a program or library over libd3q19 (the HIP runtime it
    needs is resolved through /opt/rocm/lib at link time)."""
    return [f"-L{LIB_DIR}", "-ld3q19", f"-L{ROCM_LIB}", f"-Wl,-rpath-link,{ROCM_LIB}", f"-Wl,-rpath,{rpath}"]


def build_cpp_me(verbose: bool = False) -> Path:
    """libd3q19_me.so (the moment-encoded solver's C++ interface over the C
    ABI) and its capture tool bin/lbm_me_capture."""
    common = ["-std=c++17", "-O2", "-Wall", "-Wextra", f"-I{INCLUDE}", f"-I{CPP_DIR}"]
    tmp = ME_LIB.with_suffix(".so.tmp")
    cmd = [cxx(), *common, "-shared", "-fPIC", str(CPP_DIR / "lbm_me.cpp"), str(CPP_DIR / "capture_npz.cpp"),
           "-o", str(tmp), *cpp_link_flags("$ORIGIN")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, ME_LIB)
    BIN_DIR.mkdir(exist_ok=True)
    tool = BIN_DIR / "lbm_me_capture"
    cmd = [cxx(), *common, str(CPP_DIR / "lbm_me_capture_main.cpp"), "-o", str(tool), f"-L{LIB_DIR}", "-ld3q19_me",
           *cpp_link_flags("$ORIGIN/../lib")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return ME_LIB


def build(verbose: bool = False) -> Path:
    LIB_DIR.mkdir(exist_ok=True)
    srcs = [str(CSRC / s) for s in SOURCES if (CSRC / s).exists()]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", str(tmp), *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_node_addon(verbose)
    build_cpp_me(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(f"built {LIB}", file=sys.stderr)
