"""ctypes binding of libd3q19.so (the C ABI declared in include/d3q19.h).

The library is the only compute path: if it is not built, or no GPU is
present, calls raise -- there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent.parent / "lib" / "libd3q19.so"

D3Q19_OK, D3Q19_EINVAL, D3Q19_ENOMEM, D3Q19_EHIP, D3Q19_ESTATE = 0, 1, 2, 3, 4
D3Q19_F64, D3Q19_F32, D3Q19_F16, D3Q19_BF16, D3Q19_ME16, D3Q19_BF16SR, D3Q19_ME16_REF = 0, 1, 2, 3, 4, 5, 6
BC_PERIODIC, BC_YWALLS, BC_HALFWAY = 0, 1, 2
FORCE_NONE, FORCE_UNIFORM, FORCE_FIELD = 0, 1, 2

STORAGE = {"f64": D3Q19_F64, "f32": D3Q19_F32, "f16": D3Q19_F16, "bf16": D3Q19_BF16, "bf16sr": D3Q19_BF16SR,
           "me16": D3Q19_ME16, "me16ref": D3Q19_ME16_REF}
BOUNDARY = {"periodic": BC_PERIODIC, "ywalls": BC_YWALLS, "halfway": BC_HALFWAY}
FACES = {"-x": 0, "+x": 1, "-y": 2, "+y": 3, "-z": 4, "+z": 5}

_STATUS_NAMES = {1: "EINVAL", 2: "ENOMEM", 3: "EHIP", 4: "ESTATE"}

# Every symbol include/d3q19.h declares.
EXPORTED = (
    "d3q19_last_error", "d3q19_version", "d3q19_config_default", "d3q19_create",
    "d3q19_destroy", "d3q19_set_stream", "d3q19_set_populations", "d3q19_set_equilibrium",
    "d3q19_set_force_field", "d3q19_set_solid", "d3q19_set_brinkman", "d3q19_step",
    "d3q19_step_moments", "d3q19_read_populations", "d3q19_read_moments",
    "d3q19_diagnostics", "d3q19_enstrophy", "d3q19_synchronize", "d3q19_step_count",
    "d3q19_device_bytes", "d3q19_bgk_step", "d3q19_stream", "d3q19_bounce_back_y_walls",
    "d3q19_feq_field", "d3q19_density_field", "d3q19_momentum_field",
    "d3q19_macroscopic_velocity", "d3q19_adjoint_options_default", "d3q19_adjoint_create",
    "d3q19_adjoint_destroy",
    "d3q19_adjoint_set_stream", "d3q19_adjoint_forward", "d3q19_adjoint_backward",
    "d3q19_adjoint_checkpoint_every", "d3q19_adjoint_memory_bytes", "d3q19_copy_kernel",
    "d3q19_moment_basis", "d3q19_set_moment_ranges", "d3q19_moment_range", "d3q19_moment_range_stored", "d3q19_health",
    "d3q19_set_step_count", "d3q19_state_bytes", "d3q19_read_state", "d3q19_write_state",
    "d3q19_feq", "d3q19_shader_clock", "d3q19_read_moments_noalloc", "d3q19_step_launches",
)


class D3Q19Error(RuntimeError):
    """A libd3q19 call returned a non-zero status."""

    def __init__(self, fn: str, status: int, message: str):
        super().__init__(f"{fn} failed ({_STATUS_NAMES.get(status, status)}): {message}")
        self.status = status


class Config(ctypes.Structure):
    """Mirror of d3q19_config (include/d3q19.h)."""

    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("tau", ctypes.c_double),
        ("storage", ctypes.c_int32),
        ("boundary", ctypes.c_int32),
        ("wall_velocity_top", ctypes.c_double * 3),
        ("wall_velocity_bottom", ctypes.c_double * 3),
        ("wall_faces", ctypes.c_uint32),
        ("face_velocity", (ctypes.c_double * 3) * 6),
        ("force_mode", ctypes.c_int32),
        ("force", ctypes.c_double * 3),
        ("device", ctypes.c_int32),
    ]


class AdjointOptions(ctypes.Structure):
    """Mirror of d3q19_adjoint_options (include/d3q19.h)."""

    _fields_ = [
        ("steps", ctypes.c_int32),
        ("checkpoint_every", ctypes.c_int32),
        ("checkpoint_storage", ctypes.c_int32),
        ("brinkman", ctypes.c_int32),
        ("memory_budget", ctypes.c_uint64),
    ]


_lib = None


def load() -> ctypes.CDLL:
    """Load libd3q19.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python bit-physics_amd/build.py` "
            "(the HIP library is the only compute path)"
        )
    # torch ships its own HIP runtime; load it first so that libd3q19 binds to
    # the same libamdhip64 instead of bringing a second runtime into the process
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(str(LIB_PATH))
    P, i32, i64, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    PP = ctypes.POINTER(ctypes.c_void_p)
    sigs = {
        "d3q19_last_error": (ctypes.c_char_p, []),
        "d3q19_version": (ctypes.c_char_p, []),
        "d3q19_config_default": (None, [ctypes.POINTER(Config)]),
        "d3q19_create": (ctypes.c_int, [ctypes.POINTER(Config), P, PP]),
        "d3q19_destroy": (ctypes.c_int, [P]),
        "d3q19_set_stream": (ctypes.c_int, [P, P]),
        "d3q19_set_populations": (ctypes.c_int, [P, P, i32, i32]),
        "d3q19_set_equilibrium": (ctypes.c_int, [P, P, P, i32, i32]),
        "d3q19_set_force_field": (ctypes.c_int, [P, P, i32, i32]),
        "d3q19_set_solid": (ctypes.c_int, [P, P, i32]),
        "d3q19_set_brinkman": (ctypes.c_int, [P, P, i32, i32]),
        "d3q19_step": (ctypes.c_int, [P, i32]),
        "d3q19_step_moments": (ctypes.c_int, [P, i32, P, P]),
        "d3q19_read_populations": (ctypes.c_int, [P, P, i32, i32]),
        "d3q19_read_moments": (ctypes.c_int, [P, P, P, i32, i32]),
        "d3q19_read_moments_noalloc": (ctypes.c_int, [P, P, P, i32]),
        "d3q19_diagnostics": (ctypes.c_int, [P, P]),
        "d3q19_enstrophy": (ctypes.c_int, [P, P]),
        "d3q19_synchronize": (ctypes.c_int, [P]),
        "d3q19_step_count": (i64, [P]),
        "d3q19_step_launches": (i32, [P]),
        "d3q19_device_bytes": (ctypes.c_size_t, [P]),
        "d3q19_bgk_step": (ctypes.c_int, [P, P, P, i32, i32, i32, f64, i32, P]),
        "d3q19_stream": (ctypes.c_int, [P, P, i32, i32, i32, i32, P]),
        "d3q19_bounce_back_y_walls": (ctypes.c_int, [P, P, P, P, i32, i32, i32, i32, P]),
        "d3q19_feq_field": (ctypes.c_int, [P, P, P, i64, i32, P]),
        "d3q19_feq": (ctypes.c_int, [P, P, P, i64, i32, P]),
        "d3q19_shader_clock": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_double), P]),
        "d3q19_density_field": (ctypes.c_int, [P, P, i64, i32, P]),
        "d3q19_momentum_field": (ctypes.c_int, [P, P, i64, i32, P]),
        "d3q19_macroscopic_velocity": (ctypes.c_int, [P, P, P, i64, i32, P]),
        "d3q19_adjoint_options_default": (None, [ctypes.POINTER(AdjointOptions)]),
        "d3q19_adjoint_create": (ctypes.c_int, [ctypes.POINTER(Config), ctypes.POINTER(AdjointOptions), P, PP]),
        "d3q19_adjoint_destroy": (ctypes.c_int, [P]),
        "d3q19_adjoint_set_stream": (ctypes.c_int, [P, P]),
        "d3q19_adjoint_forward": (ctypes.c_int, [P, P, P, P, P, i32, i32]),
        "d3q19_adjoint_backward": (ctypes.c_int, [P, P, P, P, P, i32, i32]),
        "d3q19_adjoint_checkpoint_every": (i32, [P]),
        "d3q19_adjoint_memory_bytes": (ctypes.c_size_t, [P]),
        "d3q19_copy_kernel": (ctypes.c_int, [P, P, ctypes.c_size_t, P]),
        "d3q19_moment_basis": (ctypes.c_int, [P, P]),
        "d3q19_set_moment_ranges": (ctypes.c_int, [P, P, P]),
        "d3q19_moment_range": (ctypes.c_int, [P, P, P]),
        "d3q19_moment_range_stored": (ctypes.c_int, [P, P, P]),
        "d3q19_health": (ctypes.c_int, [P, P]),
        "d3q19_set_step_count": (ctypes.c_int, [P, ctypes.c_int64]),
        "d3q19_state_bytes": (ctypes.c_size_t, [P]),
        "d3q19_read_state": (ctypes.c_int, [P, P, P, P]),
        "d3q19_write_state": (ctypes.c_int, [P, P, ctypes.c_int32, ctypes.c_int64]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(fn: str, status: int) -> None:
    if status != D3Q19_OK:
        msg = load().d3q19_last_error()
        raise D3Q19Error(fn, status, msg.decode() if msg else "")


def call(fn: str, *args) -> None:
    check(fn, getattr(load(), fn)(*args))
