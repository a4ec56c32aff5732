"""FP64 CPU oracle for the D3Q19 BGK step -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, and only as the checker.  The product
(``bit-physics_amd/``) never imports it.

It wraps ``oracle/d3q19_oracle.c`` (built to ``oracle/_build/``) with NumPy
in/out functions named after the reference API of
``packages/lattice-boltzmann-d3q19/lattice_boltzmann_d3q19/reference/__init__.py:43-71``.
Every array is float64, C-contiguous, direction-major ``(19, Nx, Ny, Nz)``.

Pinning (what checks each function; see tests/test_oracle_*.py; each
check of the additions also rejects a planted mistake, see planted()):
  feq, feq_field          golden table d3q19-equilibrium.json (abs 1e-15),
                          sum(W)==1, feq(1,0)==W, moment identities (PBT)
  density/momentum_field  closed-form moments of feq, brute-force sums
  stream                  np.roll (library routine) on random fields
  bgk_step                rest-state fixed point, exact mass/momentum
                          conservation, MMS ladder 4.224e-3/3.033e-3/1.543e-4
  bounce_back_y_walls     reference golden vectors (tests/golden), the
                          published Poiseuille/Couette maxima
  run_halfway             Poiseuille parabola at the exact-wall tau,
                          Couette linear profile, mass conservation; solid
                          cells: mass conservation in a box of solid blocks,
                          the parabola of a channel between solid rows (walls
                          at y = 1/2 and ny - 3/2), bit-identity with face
                          walls (tests/test_oracle_pins.py)
  diagnostics             NumPy sums written from the definitions (mass,
                          momentum, KE = 0.5 sum rho |u|^2 with the
                          Guo-shifted u), with and without force and solids
                          (tests/test_oracle_pins.py)
  inverse_grad            central finite differences of inverse_loss
  me_basis, me_moments    M M^-1 = I, rows 0..3 the density and momentum
                          moments (tests/test_oracle_moment_codec.py)
  me_encode, me_decode    round trip within the closed-form bound
                          (hi - lo) / 2 / 65535
  run_me                  mass within the quantization budget; the published
                          quantized-Poiseuille measurement against its FP64
                          parent (lbm-quantized tolerance)
"""

from __future__ import annotations

import contextlib
import ctypes
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "_build" / "libd3q19_oracle.so"
_lib = None

Q = 19


def build() -> Path:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-B", "-C", str(_HERE)], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        build()
    _lib = _bind(ctypes.CDLL(str(_LIB_PATH)))
    return _lib


PLANTS = ("KE_NO_HALF", "KE_UNIT_RHO", "KE_NO_GUO", "SOLID_SAME_DIR", "SOLID_WRONG_SIDE", "ME_FIELD_FEQ")


@contextlib.contextmanager
def planted(name: str):
    """Run the oracle's functions on a test build that carries one deliberate
    mistake (``-DORC_PLANT_<name>``, listed in d3q19_oracle.c), so a test can
    show that its independent check rejects that mistake.  Restores the
    correct build on exit."""
    global _lib
    if name not in PLANTS:
        raise ValueError(f"unknown planted mistake {name!r}; known: {PLANTS}")
    target = f"_build/libd3q19_oracle_plant_{name}.so"
    subprocess.run(["make", "-s", "-C", str(_HERE), target], check=True)
    saved = _load()
    _lib = _bind(ctypes.CDLL(str(_HERE / target)))
    try:
        yield
    finally:
        _lib = saved


def _bind(lib):
    P = ctypes.c_void_p
    i32, i64, f64, u32 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_uint32
    sigs = {
        "orc_constants": (None, [P, P, P, P]),
        "orc_feq_point": (None, [f64, f64, f64, f64, P]),
        "orc_feq_field": (None, [i64, P, P, P]),
        "orc_density_field": (None, [i64, P, P]),
        "orc_momentum_field": (None, [i64, P, P]),
        "orc_macroscopic_velocity": (None, [i64, P, P, P]),
        "orc_collide": (None, [i64, P, P, f64, P]),
        "orc_me_collide": (None, [i64, P, P, f64, P]),
        "orc_stream": (None, [i32, i32, i32, P, P]),
        "orc_bgk_step": (None, [i32, i32, i32, P, P, f64, P, P]),
        "orc_bounce_back_y_walls": (None, [i32, i32, i32, P, P, P, P]),
        "orc_run_reference": (None, [i32, i32, i32, P, f64, P, i32, P, P, i32, P]),
        "orc_stream_halfway": (None, [i32, i32, i32, P, P, u32, P, P]),
        "orc_run_halfway": (None, [i32, i32, i32, P, f64, P, P, P, u32, P, P, i32, P]),
        "orc_diagnostics": (None, [i64, P, P, P, P]),
        "orc_inverse_loss": (f64, [i32, i32, i32, P, P, P, f64, i32]),
        "orc_inverse_grad": (f64, [i32, i32, i32, P, P, P, f64, i32, P, P]),
        "orc_me_basis": (f64, [P, P]),
        "orc_me_encode": (None, [i64, P, P, P, P]),
        "orc_me_decode": (None, [i64, P, P, P, P]),
        "orc_run_me": (None, [i32, i32, i32, P, f64, P, i32, P, P, i32, P, P, P, P]),
        "orc_run_me_codes": (None, [i32, i32, i32, P, f64, P, i32, P, P, i32, P, P, P, P]),
        "orc_max_threads": (i32, []),
        "orc_set_threads": (None, [i32]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def _p(a):
    # data_as keeps a reference to the array: a temporary such as _f64(view)
    # stays alive for as long as the pointer (the call's argument) does
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def constants():
    lib = _load()
    c = np.zeros((19, 3), np.int32)
    w = np.zeros(19, np.float64)
    opp = np.zeros(19, np.int32)
    cs2 = ctypes.c_double()
    lib.orc_constants(_p(c), _p(w), _p(opp), ctypes.byref(cs2))
    return c, w, opp, cs2.value


def feq(rho: float, u) -> list[float]:
    out = np.zeros(19, np.float64)
    _load().orc_feq_point(float(rho), float(u[0]), float(u[1]), float(u[2]), _p(out))
    return out.tolist()


def feq_field(rho, u):
    rho = _f64(rho)
    u = _f64(u)
    out = np.empty((19,) + rho.shape, np.float64)
    _load().orc_feq_field(rho.size, _p(rho), _p(u), _p(out))
    return out


def density_field(f):
    f = _f64(f)
    out = np.empty(f.shape[1:], np.float64)
    _load().orc_density_field(out.size, _p(f), _p(out))
    return out


def momentum_field(f):
    f = _f64(f)
    out = np.empty((3,) + f.shape[1:], np.float64)
    _load().orc_momentum_field(out[0].size, _p(f), _p(out))
    return out


def macroscopic_velocity(f, force_lattice=None):
    f = _f64(f)
    F = None if force_lattice is None else _f64(np.broadcast_to(force_lattice, (3,) + f.shape[1:]))
    out = np.empty((3,) + f.shape[1:], np.float64)
    _load().orc_macroscopic_velocity(out[0].size, _p(f), _p(F), _p(out))
    return out


def collide(f, tau, force_lattice=None):
    f = _f64(f)
    F = None if force_lattice is None else _f64(np.broadcast_to(force_lattice, (3,) + f.shape[1:]))
    out = np.empty_like(f)
    _load().orc_collide(f[0].size, _p(f), _p(out), float(tau), _p(F))
    return out


def me_collide(f, tau, force_lattice=None):
    """The moment-encoded variant's collision (ME/shaders/lbm_collide.comp):
    collide() with the equilibrium in the division form of the point function."""
    f = _f64(f)
    F = None if force_lattice is None else _f64(np.broadcast_to(force_lattice, (3,) + f.shape[1:]))
    out = np.empty_like(f)
    _load().orc_me_collide(f[0].size, _p(f), _p(out), float(tau), _p(F))
    return out


def stream(f_post):
    f_post = _f64(f_post)
    out = np.empty_like(f_post)
    nx, ny, nz = f_post.shape[1:]
    _load().orc_stream(nx, ny, nz, _p(f_post), _p(out))
    return out


def bgk_step(f, tau, force_lattice=None):
    f = _f64(f)
    nx, ny, nz = f.shape[1:]
    F = None if force_lattice is None else _f64(np.broadcast_to(force_lattice, (3, nx, ny, nz)))
    out = np.empty_like(f)
    scratch = np.empty_like(f)
    _load().orc_bgk_step(nx, ny, nz, _p(f), _p(out), float(tau), _p(F), _p(scratch))
    return out


def apply_bounce_back_y_walls(f, wall_velocity_top=(0.0, 0.0, 0.0),
                              wall_velocity_bottom=(0.0, 0.0, 0.0)):
    f = _f64(f)
    nx, ny, nz = f.shape[1:]
    top = _f64(wall_velocity_top)
    bot = _f64(wall_velocity_bottom)
    out = np.empty_like(f)
    _load().orc_bounce_back_y_walls(nx, ny, nz, _p(f), _p(out), _p(top), _p(bot))
    return out


def run_reference(f, tau, steps, force_lattice=None, ywalls=False,
                  wall_velocity_top=(0.0, 0.0, 0.0), wall_velocity_bottom=(0.0, 0.0, 0.0)):
    """NP/sim.py:312-326 loop: bgk_step then (optionally) the y-wall bounce-back."""
    f = _f64(f).copy()
    nx, ny, nz = f.shape[1:]
    F = None if force_lattice is None else _f64(np.broadcast_to(force_lattice, (3, nx, ny, nz)))
    top = _f64(wall_velocity_top)
    bot = _f64(wall_velocity_bottom)
    work = np.empty((2,) + f.shape, np.float64)
    _load().orc_run_reference(nx, ny, nz, _p(f), float(tau), _p(F), int(bool(ywalls)),
                              _p(top), _p(bot), int(steps), _p(work))
    return f


FACE_NAMES = ("-x", "+x", "-y", "+y", "-z", "+z")


def run_halfway(f, tau, steps, wall_mask=0, face_velocity=None, solid=None,
                force_field=None, force_uniform=None, brinkman=None):
    """Half-way bounce-back engine mode (no reference counterpart)."""
    f = _f64(f).copy()
    nx, ny, nz = f.shape[1:]
    fu = np.zeros((6, 3)) if face_velocity is None else face_velocity
    fu = _f64(fu)
    sol = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8)
    F = None if force_field is None else _f64(force_field)
    Fu = None if force_uniform is None else _f64(force_uniform)
    B = None if brinkman is None else _f64(brinkman)
    work = np.empty_like(f)
    _load().orc_run_halfway(nx, ny, nz, _p(f), float(tau), _p(F), _p(Fu), _p(B),
                            int(wall_mask), _p(fu), _p(sol), int(steps), _p(work))
    return f


def diagnostics(f, force_uniform=None, solid=None):
    """[mass, px, py, pz, kinetic energy] in FP64 (sequential sums)."""
    f = _f64(f)
    out = np.zeros(5, np.float64)
    Fu = None if force_uniform is None else _f64(force_uniform)
    sol = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8)
    _load().orc_diagnostics(f[0].size, _p(f), _p(Fu), _p(sol), _p(out))
    return out


def inverse_loss(u0, F, u_tgt, tau, T):
    u0, F, u_tgt = _f64(u0), _f64(F), _f64(u_tgt)
    nx, ny, nz = u0.shape[1:]
    return _load().orc_inverse_loss(nx, ny, nz, _p(u0), _p(F), _p(u_tgt), float(tau), int(T))


def inverse_grad(u0, F, u_tgt, tau, T):
    u0, F, u_tgt = _f64(u0), _f64(F), _f64(u_tgt)
    nx, ny, nz = u0.shape[1:]
    g_u0 = np.zeros_like(u0)
    g_F = np.zeros_like(F)
    L = _load().orc_inverse_grad(nx, ny, nz, _p(u0), _p(F), _p(u_tgt), float(tau), int(T),
                                 _p(g_u0), _p(g_F))
    return L, g_u0, g_F


def me_basis():
    """(M, M^-1, max |M M^-1 - I|) of the moment-encoded variant's basis."""
    m = np.empty((19, 19), np.float64)
    minv = np.empty((19, 19), np.float64)
    resid = _load().orc_me_basis(_p(m), _p(minv))
    return m, minv, resid


def me_encode(f, lo, hi):
    """16-bit codes (19, Nx, Ny, Nz), moment-major, of populations f."""
    f = _f64(f)
    lo, hi = _f64(lo), _f64(hi)
    codes = np.empty(f.shape, np.uint16)
    _load().orc_me_encode(f[0].size, _p(f), _p(lo), _p(hi), _p(codes))
    return codes


def me_decode(codes, lo, hi):
    codes = np.ascontiguousarray(codes, dtype=np.uint16)
    lo, hi = _f64(lo), _f64(hi)
    f = np.empty(codes.shape, np.float64)
    _load().orc_me_decode(codes[0].size, _p(codes), _p(lo), _p(hi), _p(f))
    return f


def run_me(f, tau, steps, lo, hi, force_lattice=None, ywalls=False,
           wall_velocity_top=(0.0, 0.0, 0.0), wall_velocity_bottom=(0.0, 0.0, 0.0), return_codes=False,
           codes=None):
    """Quantized trajectory of the moment-encoded variant (ME/src/lbm_me.cpp:
    376-409): encode f, then per step decode, collide (me_collide), stream,
    y-walls, encode; returns the decoded final state (and, with
    ``return_codes``, the final codes (19, Nx, Ny, Nz) uint16).

    ``codes`` (uint16, the shape of f) continues a trajectory from its codes
    instead of encoding f (f may then be None).  Decoding then re-encoding is
    not the identity for a moment whose range is narrower than FP64 resolves
    around its value, so a run restarted from decoded populations is another
    trajectory than the one continued from its codes."""
    if codes is None:
        f = _f64(f).copy()
        shape = f.shape
    else:
        shape = codes.shape
        f = np.empty(shape, np.float64)
    nx, ny, nz = shape[1:]
    F = None if force_lattice is None else _f64(np.broadcast_to(force_lattice, (3, nx, ny, nz)))
    work = np.empty((2,) + shape, np.float64)
    args = (float(tau), _p(F), int(bool(ywalls)), _p(_f64(wall_velocity_top)), _p(_f64(wall_velocity_bottom)),
            int(steps), _p(_f64(lo)), _p(_f64(hi)))
    if codes is None:
        c = np.empty(shape, np.uint16)
        _load().orc_run_me(nx, ny, nz, _p(f), *args, _p(c), _p(work))
    else:
        c = np.array(codes, dtype=np.uint16, order="C", copy=True)
        _load().orc_run_me_codes(nx, ny, nz, _p(c), *args, _p(f), _p(work))
    return (f, c) if return_codes else f


def me_moments(f):
    """(19, Nx, Ny, Nz) moments M f of populations f (FP64, sums in basis order)."""
    m, _, _ = me_basis()
    f = _f64(f)
    return np.tensordot(m, f, axes=(1, 0))


def num_threads() -> int:
    """Threads the oracle's parallel loops use now."""
    return int(_load().orc_max_threads())


def set_num_threads(n: int) -> None:
    _load().orc_set_threads(int(n))
