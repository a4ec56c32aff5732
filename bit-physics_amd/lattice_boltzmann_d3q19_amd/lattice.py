"""Stateful lattice: the step / read-back path as one object over the C ABI.

Replaces the runner loop of the reference
(packages/lattice-boltzmann-d3q19/lattice_boltzmann_d3q19/sim.py:312-326 and
:350-364)::

    f = bgk_step(f, tau, force_lattice=F)
    f = apply_bounce_back_y_walls(f, wall_velocity_top=..., wall_velocity_bottom=...)

with ``Lattice.step(n)``: one fused pull-stream / bounce-back / BGK+Guo
kernel per step on the GPU.  ``populations()`` returns the reference's
state f^n (pre-collision), ``moments()`` returns ``density_field(f^n)``
and ``macroscopic_velocity(f^n, F)``.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _device as dev
from ._lib import (
    BOUNDARY,
    D3Q19_F32,
    D3Q19_F64,
    FACES,
    FORCE_FIELD,
    FORCE_NONE,
    FORCE_UNIFORM,
    STORAGE,
    Config,
    call,
    load,
)


def _host_or_device(a, dtype_np):
    """(pointer, dtype code, on_device, keepalive) for a NumPy array or a device tensor."""
    if dev.is_tensor(a):
        t = dev.torch()
        if not a.is_cuda:
            a = a.cpu().numpy()
        else:
            if a.dtype not in (t.float32, t.float64):
                a = a.to(t.float64)
            a = a.contiguous()
            code = D3Q19_F64 if a.dtype == t.float64 else D3Q19_F32
            return ctypes.c_void_p(a.data_ptr()), code, 1, a
    arr = np.ascontiguousarray(a, dtype=dtype_np)
    code = D3Q19_F64 if arr.dtype == np.float64 else D3Q19_F32
    return ctypes.c_void_p(arr.ctypes.data), code, 0, arr


class Lattice:
    """A D3Q19 lattice resident on one GPU.

    Parameters
    ----------
    shape : (Nx, Ny, Nz)
    tau : BGK relaxation time (> 0.5)
    storage : "f32" (default) | "f64" | "f16" | "bf16" | "bf16sr" | "me16" | "me16ref" --
        population storage; 16-bit formats store f_i - w_i ("bf16sr":
        stochastically rounded) or the 16-bit codes of the 19 moments of
        f - w ("me16": the stored post-collision state, one fused pass per
        step; ranges from moment_range(stored=True) of an FP64 run);
        collision stays FP32.  "me16ref": moment codes of f^n stepped in the
        moment-encoded reference variant's FP64 arithmetic (periodic or
        y-walls; no reductions).
    boundary : "periodic" | "ywalls" (the reference's apply_bounce_back_y_walls)
        | "halfway" (textbook half-way bounce-back on `wall_faces`, solids).
    wall_velocity_top / wall_velocity_bottom : ywalls mode wall velocities.
    wall_faces : iterable of "-x","+x","-y","+y","-z","+z" (halfway mode).
    face_velocity : {face: (ux, uy, uz)} moving walls (halfway mode).
    force : None, a uniform (Fx, Fy, Fz), or a (3, Nx, Ny, Nz) field (Guo).
    """

    def __init__(self, shape, tau, *, storage="f32", boundary="periodic",
                 wall_velocity_top=(0.0, 0.0, 0.0), wall_velocity_bottom=(0.0, 0.0, 0.0),
                 wall_faces=(), face_velocity=None, force=None, device=0, stream=None):
        lib = load()
        nx, ny, nz = (int(s) for s in shape)
        self.shape = (nx, ny, nz)
        self.ncell = nx * ny * nz
        self.tau = float(tau)
        self.storage = storage
        self.boundary = boundary
        self.device = int(device)
        cfg = Config()
        lib.d3q19_config_default(ctypes.byref(cfg))
        cfg.nx, cfg.ny, cfg.nz = nx, ny, nz
        cfg.tau = self.tau
        try:
            cfg.storage = STORAGE[storage]
            cfg.boundary = BOUNDARY[boundary]
        except KeyError as e:
            raise ValueError(f"unknown storage/boundary {e}") from None
        for a in range(3):
            cfg.wall_velocity_top[a] = float(wall_velocity_top[a])
            cfg.wall_velocity_bottom[a] = float(wall_velocity_bottom[a])
        mask = 0
        for face in wall_faces:
            mask |= 1 << FACES[face]
        cfg.wall_faces = mask
        for face, vel in (face_velocity or {}).items():
            for a in range(3):
                cfg.face_velocity[FACES[face]][a] = float(vel[a])
        field = None  # copied into the library below; no reference is kept
        self._me_ranges = None
        if force is None:
            cfg.force_mode = FORCE_NONE
        elif np.ndim(force) == 1:
            cfg.force_mode = FORCE_UNIFORM
            for a in range(3):
                cfg.force[a] = float(force[a])
        else:
            cfg.force_mode = FORCE_FIELD
            field = force
        cfg.device = self.device
        self._cfg = cfg
        self._stream = stream
        h = ctypes.c_void_p()
        call("d3q19_create", ctypes.byref(cfg), self._stream_arg(), ctypes.byref(h))
        self._h = h
        if field is not None:
            self.set_force_field(field)

    # ---- plumbing -------------------------------------------------------
    def _stream_arg(self):
        if self._stream is not None:
            return ctypes.c_void_p(int(self._stream))
        return dev.stream_ptr(self.device)

    def _sync_stream(self):
        call("d3q19_set_stream", self._h, self._stream_arg())

    def _check_out(self, t, name, shape, dtypes):
        """Validate a device tensor that a kernel will write through its raw
        pointer: a CUDA tensor on the lattice's device, contiguous, of exactly
        `shape` and one of `dtypes`.  Raises ValueError before any launch."""
        if t is None:
            return
        if not dev.is_tensor(t):
            raise ValueError(f"{name} must be a torch device tensor, got {type(t).__name__}")
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
        if t.dtype not in dtypes:
            raise ValueError(f"{name} must have dtype in {[str(d) for d in dtypes]}, got {t.dtype}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
        if not t.is_cuda or t.device.index != self.device:
            raise ValueError(f"{name} must live on cuda:{self.device}, got {t.device}")

    def _copied(self, on_dev):
        """A device input (or its converted copy) is read by work queued on the
        lattice's stream.  On torch's current stream, torch's allocator orders
        any reuse of the tensor's memory after that work; on a caller's own
        stream it does not, so wait for the copy there."""
        if on_dev and self._stream is not None:
            call("d3q19_synchronize", self._h)

    def close(self):
        if getattr(self, "_h", None):
            load().d3q19_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- state ----------------------------------------------------------
    def set_equilibrium(self, rho=None, u=None):
        """State <- feq_field(rho, u) (equilibrium.py:71-98); rho None means 1, u None means 0."""
        self._sync_stream()
        grid = self.shape
        if u is None:
            u = np.zeros((3,) + grid, np.float64)
        if tuple(u.shape) != (3,) + grid:
            raise ValueError(f"u must have shape {(3,) + grid}")
        pu, code, on_dev, keep_u = _host_or_device(u, np.float64)
        prho = ctypes.c_void_p(None)
        keep_r = None
        if rho is not None:
            if tuple(rho.shape) != grid:
                raise ValueError(f"rho must have shape {grid}")
            if dev.is_tensor(keep_u):
                keep_r = rho.to(dtype=keep_u.dtype, device=keep_u.device).contiguous()
                prho = ctypes.c_void_p(keep_r.data_ptr())
            else:
                keep_r = np.ascontiguousarray(rho, dtype=keep_u.dtype)
                prho = ctypes.c_void_p(keep_r.ctypes.data)
        call("d3q19_set_equilibrium", self._h, prho, pu, code, on_dev)
        return self

    def set_populations(self, f):
        """State <- f (pre-collision populations, shape (19, Nx, Ny, Nz))."""
        if tuple(f.shape) != (19,) + self.shape:
            raise ValueError(f"populations must have shape {(19,) + self.shape}")
        self._sync_stream()
        p, code, on_dev, keep = _host_or_device(f, np.float64)
        call("d3q19_set_populations", self._h, p, code, on_dev)
        self._copied(on_dev)
        return self

    def set_force_field(self, force):
        """Per-cell Guo body force (3, Nx, Ny, Nz) used from the next collision on."""
        if tuple(force.shape) != (3,) + self.shape:
            raise ValueError(f"force must have shape {(3,) + self.shape}")
        if self._cfg.force_mode != FORCE_FIELD:
            raise ValueError("lattice was not created with a force field")
        self._sync_stream()
        p, code, on_dev, keep = _host_or_device(force, np.float64)
        call("d3q19_set_force_field", self._h, p, code, on_dev)
        self._copied(on_dev)
        return self

    def set_solid(self, solid):
        """Solid cells (Nx, Ny, Nz) bool/uint8 for the half-way mode; None clears."""
        self._sync_stream()
        if solid is None:
            call("d3q19_set_solid", self._h, ctypes.c_void_p(None), 0)
            return self
        if dev.is_tensor(solid):
            t = dev.torch()
            s = solid.to(dtype=t.uint8, device=f"cuda:{self.device}").contiguous()
            call("d3q19_set_solid", self._h, ctypes.c_void_p(s.data_ptr()), 1)
        else:
            s = np.ascontiguousarray(solid, dtype=np.uint8)
            call("d3q19_set_solid", self._h, ctypes.c_void_p(s.ctypes.data), 0)
        return self

    def set_brinkman(self, s):
        """Brinkman solid fraction in [0, 1] (Nx, Ny, Nz); None clears."""
        self._sync_stream()
        if s is None:
            call("d3q19_set_brinkman", self._h, ctypes.c_void_p(None), D3Q19_F64, 0)
            return self
        p, code, on_dev, keep = _host_or_device(s, np.float64)
        call("d3q19_set_brinkman", self._h, p, code, on_dev)
        self._copied(on_dev)
        return self

    # ---- stepping -------------------------------------------------------
    def step(self, n: int = 1):
        self._sync_stream()
        call("d3q19_step", self._h, int(n))
        return self

    def step_moments(self, n: int, rho_out=None, u_out=None):
        """Advance n steps; the last step's pass also writes rho/u (FP32 device
        tensors) of the state that step started from."""
        if rho_out is not None or u_out is not None:
            t = dev.torch()
            self._check_out(rho_out, "rho_out", self.shape, (t.float32,))
            self._check_out(u_out, "u_out", (3,) + self.shape, (t.float32,))
        self._sync_stream()
        call("d3q19_step_moments", self._h, int(n), dev.ptr(rho_out), dev.ptr(u_out))
        return self

    @property
    def step_count(self) -> int:
        return int(load().d3q19_step_count(self._h))

    @property
    def step_launches(self) -> int:
        """Kernel launches one step makes (the fused step's plane chunks)."""
        return int(load().d3q19_step_launches(self._h))

    @property
    def device_bytes(self) -> int:
        return int(load().d3q19_device_bytes(self._h))

    def synchronize(self):
        call("d3q19_synchronize", self._h)

    # ---- read-back ------------------------------------------------------
    def populations(self, dtype=np.float64):
        """f^n, the reference's pre-collision state, as a NumPy array."""
        self._sync_stream()
        out = np.empty((19,) + self.shape, dtype=dtype)
        code = D3Q19_F64 if out.dtype == np.float64 else D3Q19_F32
        call("d3q19_read_populations", self._h, ctypes.c_void_p(out.ctypes.data), code, 0)
        return out

    def populations_into(self, out):
        """Write f^n into a (19, Nx, Ny, Nz) float32/float64 device tensor."""
        t = dev.torch()
        if out is None:
            raise ValueError("out must be a (19, Nx, Ny, Nz) device tensor")
        self._check_out(out, "out", (19,) + self.shape, (t.float32, t.float64))
        self._sync_stream()
        code = D3Q19_F64 if out.dtype == t.float64 else D3Q19_F32
        call("d3q19_read_populations", self._h, dev.ptr(out), code, 1)
        return out

    def moments(self, dtype=np.float64):
        """(rho, u) of f^n: density_field and macroscopic_velocity(f, F)."""
        self._sync_stream()
        rho = np.empty(self.shape, dtype=dtype)
        u = np.empty((3,) + self.shape, dtype=dtype)
        code = D3Q19_F64 if rho.dtype == np.float64 else D3Q19_F32
        call("d3q19_read_moments", self._h, ctypes.c_void_p(rho.ctypes.data),
             ctypes.c_void_p(u.ctypes.data), code, 0)
        return rho, u

    def moments_into(self, rho=None, u=None):
        """Write rho (Nx, Ny, Nz) and/or u (3, Nx, Ny, Nz) of f^n into device
        tensors of one dtype (float32 or float64); either may be None."""
        if rho is None and u is None:
            return rho, u
        t = dev.torch()
        ref = rho if rho is not None else u
        if getattr(ref, "dtype", None) not in (t.float32, t.float64):
            raise ValueError(f"rho/u must be float32 or float64 device tensors, got {getattr(ref, 'dtype', ref)}")
        self._check_out(rho, "rho", self.shape, (ref.dtype,))
        self._check_out(u, "u", (3,) + self.shape, (ref.dtype,))
        self._sync_stream()
        code = D3Q19_F64 if ref.dtype == t.float64 else D3Q19_F32
        call("d3q19_read_moments", self._h, dev.ptr(rho), dev.ptr(u), code, 1)
        return rho, u

    def diagnostics(self) -> dict:
        """FP64 mass, momentum and kinetic energy of f^n."""
        self._sync_stream()
        out = (ctypes.c_double * 5)()
        call("d3q19_diagnostics", self._h, out)
        return {"mass": out[0], "momentum": (out[1], out[2], out[3]), "kinetic_energy": out[4]}

    # ---- checkpoint / restore -----------------------------------------
    def raw_state(self):
        """(bytes, local, step_count) of the stored state: the population
        buffer in its storage format and layout (uint8 array), whether it holds
        f^n in place (not yet streamed), and the step count."""
        lib = load()
        raw = np.empty(int(lib.d3q19_state_bytes(self._h)), np.uint8)
        local = ctypes.c_int32()
        steps = ctypes.c_int64()
        self._sync_stream()
        call("d3q19_read_state", self._h, ctypes.c_void_p(raw.ctypes.data), ctypes.byref(local),
             ctypes.byref(steps))
        return raw, bool(local.value), int(steps.value)

    def checkpoint(self, path):
        """Write the stored state (the population buffer in its storage format
        and layout, whether it holds f^n in place, the step count) and, for
        "me16", the moment ranges to ``path`` (.npz).  Configuration and
        fields (force field, solids, Brinkman) are not saved: restore into a
        lattice created and set up the same way."""
        raw, local, steps = self.raw_state()
        arrays = {"state": raw, "local": np.int32(local), "step_count": np.int64(steps),
                  "shape": np.asarray(self.shape, np.int64), "tau": np.float64(self.tau),
                  "storage": np.asarray(self.storage), "boundary": np.asarray(self.boundary)}
        if self._me_ranges is not None:
            arrays["moment_lo"], arrays["moment_hi"] = self._me_ranges
        with open(path, "wb") as fh:
            np.savez(fh, **arrays)
        return path

    def restore(self, path):
        """Load a checkpoint written by :meth:`checkpoint`; the continued run
        repeats the uninterrupted one bit for bit, in every storage."""
        with np.load(path, allow_pickle=False) as z:
            if tuple(z["shape"].tolist()) != self.shape or float(z["tau"]) != self.tau:
                raise ValueError("checkpoint shape / tau differ from this lattice")
            if str(z["storage"]) != self.storage or str(z["boundary"]) != self.boundary:
                raise ValueError(f"checkpoint storage/boundary {z['storage']}/{z['boundary']} differ "
                                 f"from {self.storage}/{self.boundary}")
            raw = np.ascontiguousarray(z["state"])
            if raw.nbytes != int(load().d3q19_state_bytes(self._h)):
                raise ValueError("checkpoint state size differs from this lattice")
            if "moment_lo" in z.files:
                lo, hi = z["moment_lo"], z["moment_hi"]
                if self._me_ranges is None:
                    self.set_moment_ranges(lo, hi)
                elif not (np.array_equal(lo, self._me_ranges[0]) and np.array_equal(hi, self._me_ranges[1])):
                    # the codes only mean the saved state under the ranges they were made with
                    raise ValueError("checkpoint moment ranges differ from this lattice's")
            self._sync_stream()
            call("d3q19_write_state", self._h, ctypes.c_void_p(raw.ctypes.data), int(z["local"]),
                 int(z["step_count"]))
        return self

    def health(self) -> dict:
        """Tier-1 health of f^n on the GPU (NP/sim.py:135-145): check_health
        0.0 when every population, rho and u is finite, else 1.0; the number of
        non-finite cells; rho range and largest |u| over the finite cells."""
        self._sync_stream()
        out = (ctypes.c_double * 4)()
        call("d3q19_health", self._h, out)
        return {"check_health": 0.0 if out[0] == 0 else 1.0, "nonfinite_cells": int(out[0]),
                "rho_min": out[1], "rho_max": out[2], "u_max_lat": out[3]}

    # ---- moment-encoded storage ----------------------------------------
    def set_moment_ranges(self, lo, hi):
        """Quantization ranges [lo_k, hi_k] of the 19 moments of f ("me16" /
        "me16ref" storage; before the first state is loaded)."""
        lo = np.ascontiguousarray(lo, dtype=np.float64)
        hi = np.ascontiguousarray(hi, dtype=np.float64)
        if lo.shape != (19,) or hi.shape != (19,):
            raise ValueError("lo and hi must hold 19 values")
        call("d3q19_set_moment_ranges", self._h, ctypes.c_void_p(lo.ctypes.data), ctypes.c_void_p(hi.ctypes.data))
        self._me_ranges = (lo.copy(), hi.copy())
        return self

    def moment_range(self, stored: bool = False):
        """(lo, hi): FP64 minimum and maximum over all cells of the 19 moments
        m = M f of the current state f^n, or with ``stored`` of the state as
        the lattice stores it (the post-collision populations of the last
        step; f^n after a load) -- the states an "me16" lattice encodes."""
        self._sync_stream()
        lo = np.empty(19, np.float64)
        hi = np.empty(19, np.float64)
        fn = "d3q19_moment_range_stored" if stored else "d3q19_moment_range"
        call(fn, self._h, ctypes.c_void_p(lo.ctypes.data), ctypes.c_void_p(hi.ctypes.data))
        return lo, hi

    def enstrophy(self) -> float:
        self._sync_stream()
        out = ctypes.c_double()
        call("d3q19_enstrophy", self._h, ctypes.byref(out))
        return out.value
