"""The C ABI library loads and exports every symbol include/*.h declares;
the Python mirror of d3q19_config has the C layout (CPU, no compute calls)."""

from __future__ import annotations

import ctypes
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_functions():
    names = []
    for h in HEADERS:
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(d3q19_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_headers_declare_functions():
    names = declared_functions()
    assert "d3q19_step" in names and "d3q19_bgk_step" in names
    assert len(names) >= 27


def test_library_exports_every_declared_symbol():
    from lattice_boltzmann_d3q19_amd._lib import LIB_PATH

    if not LIB_PATH.exists():
        pytest.fail(f"{LIB_PATH} not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    # the version string needs no GPU
    lib.d3q19_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.d3q19_version()


def test_python_binding_covers_header():
    from lattice_boltzmann_d3q19_amd._lib import EXPORTED

    assert set(declared_functions()) <= set(EXPORTED)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc needed to read the C layout")
def test_config_struct_layout_matches_c(tmp_path):
    from lattice_boltzmann_d3q19_amd._lib import Config

    fields = [f[0] for f in Config._fields_]
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{ROOT / "include" / "d3q19.h"}"',
             "int main(void) {", 'printf("%zu\\n", sizeof(d3q19_config));']
    lines += [f'printf("%zu\\n", offsetof(d3q19_config, {f}));' for f in fields]
    lines += ["return 0; }"]
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == ctypes.sizeof(Config)
    for f, off in zip(fields, out[1:]):
        assert getattr(Config, f).offset == off, f


def test_config_default_needs_no_gpu():
    from lattice_boltzmann_d3q19_amd._lib import Config, load

    cfg = Config()
    load().d3q19_config_default(ctypes.byref(cfg))
    assert cfg.tau == 0.6 and cfg.storage == 1 and cfg.boundary == 0


def test_create_rejects_bad_config_without_gpu():
    """Argument validation runs before any HIP call."""
    from lattice_boltzmann_d3q19_amd._lib import Config, load

    lib = load()
    cfg = Config()
    lib.d3q19_config_default(ctypes.byref(cfg))
    cfg.nx = cfg.ny = cfg.nz = 4
    h = ctypes.c_void_p()
    cfg.tau = 0.5
    assert lib.d3q19_create(ctypes.byref(cfg), None, ctypes.byref(h)) == 1  # EINVAL
    assert b"tau" in lib.d3q19_last_error()
    cfg.tau = 0.6
    cfg.storage = 9
    assert lib.d3q19_create(ctypes.byref(cfg), None, ctypes.byref(h)) == 1
    cfg.storage = 1
    cfg.nx = 0
    assert lib.d3q19_create(ctypes.byref(cfg), None, ctypes.byref(h)) == 1
    assert not h.value


def test_node_addon_loads_without_gpu():
    import shutil as _sh

    node = _sh.which("node")
    addon = ROOT / "bit-physics_amd" / "node" / "build" / "d3q19.node"
    if node is None or not addon.exists():
        pytest.skip("node or the addon is not available")
    r = subprocess.run([node, "-e", "const m=require(process.argv[1]); console.log(typeof m.Lattice, m.version())",
                        str(ROOT / "bit-physics_amd" / "node" / "index.js")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("function d3q19")


# reference/__init__.py:43-71 of packages/lattice-boltzmann-d3q19/lattice_boltzmann_d3q19
REFERENCE_SURFACE = {
    "C", "CANONICAL_COUETTE_NX", "CANONICAL_COUETTE_NY", "CANONICAL_COUETTE_STEPS",
    "CANONICAL_DESCRIPTOR_COUETTE", "CANONICAL_DESCRIPTOR_POISEUILLE", "CANONICAL_NZ",
    "CANONICAL_POISEUILLE_NX", "CANONICAL_POISEUILLE_NY", "CANONICAL_POISEUILLE_STEPS",
    "CANONICAL_SEED", "CS2", "VELOCITIES", "W", "WEIGHTS", "apply_bounce_back_y_walls", "bgk",
    "bgk_step", "density_field", "density_moment", "equilibrium", "feq", "feq_field",
    "macroscopic_velocity", "momentum_field", "momentum_moment", "stream",
}


def test_dropin_surface_matches_reference():
    """Same public names and submodule import paths as the reference package
    (its tests import ``reference.bgk.bgk_step`` and ``reference.equilibrium``)."""
    import importlib

    import numpy as np

    R = importlib.import_module("lattice_boltzmann_d3q19_amd.reference")
    assert set(R.__all__) == REFERENCE_SURFACE
    assert all(hasattr(R, n) for n in REFERENCE_SURFACE)
    bgk = importlib.import_module("lattice_boltzmann_d3q19_amd.reference.bgk")
    eq = importlib.import_module("lattice_boltzmann_d3q19_amd.reference.equilibrium")
    assert R.bgk is bgk and R.equilibrium is eq and bgk.bgk_step is R.bgk_step
    assert {"feq_field", "feq", "density_field", "momentum_field"} <= set(dir(eq))
    # constants.py: weights sum to one, c_i opposite pairs, isotropy sum w c c = cs2 I
    assert abs(sum(R.WEIGHTS) - 1.0) < 1e-15 and R.CS2 == 1.0 / 3.0
    assert np.array_equal(R.C[R.OPP], -R.C)
    assert np.allclose(np.einsum("i,ia,ib->ab", R.W, R.C, R.C), np.eye(3) / 3, rtol=0, atol=1e-15)
    assert R.CANONICAL_DESCRIPTOR_POISEUILLE == "poiseuille-64x32-seed42-step1000"


def test_output_tensor_checks_reject_bad_tensors_without_gpu():
    """Lattice's device-output paths (populations_into, moments_into,
    step_moments) hand raw pointers to kernels that write through them: the
    tensor's shape, dtype, contiguity and device are checked first, and a bad
    one raises ValueError before any launch."""
    import types

    import numpy as np
    import torch

    from lattice_boltzmann_d3q19_amd.lattice import Lattice

    fake = types.SimpleNamespace(device=0)
    check = Lattice._check_out.__get__(fake)
    shape = (19, 4, 3, 2)
    ok_dtypes = (torch.float32, torch.float64)
    with pytest.raises(ValueError, match="shape"):
        check(torch.empty((19, 4, 3, 3)), "out", shape, ok_dtypes)
    with pytest.raises(ValueError, match="dtype"):
        check(torch.empty(shape, dtype=torch.float16), "out", shape, ok_dtypes)
    with pytest.raises(ValueError, match="contiguous"):
        check(torch.empty((19, 2, 3, 4)).transpose(1, 3), "out", shape, ok_dtypes)
    with pytest.raises(ValueError, match="cuda:0"):
        check(torch.empty(shape), "out", shape, ok_dtypes)  # a host tensor
    with pytest.raises(ValueError, match="torch device tensor"):
        check(np.empty(shape), "out", shape, ok_dtypes)
    check(None, "out", shape, ok_dtypes)  # absent outputs are allowed
