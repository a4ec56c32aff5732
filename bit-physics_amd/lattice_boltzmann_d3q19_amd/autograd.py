"""torch.autograd.Function over the engine's adjoint (C ABI d3q19_adjoint_*).

    solver = DifferentiableLBM((128, 128, 128), tau=0.6, steps=2000, checkpoint_every=50)
    u_T = solver(u0, F)                 # u0, F: (3, Nx, Ny, Nz) device tensors
    loss = ((u_T - u_target) ** 2).sum(0).mean()
    loss.backward()                     # u0.grad, F.grad via the adjoint kernels

Forward: f^0 = feq(1, u0); ``steps`` BGK+Guo steps with the per-cell force F
(the reference's ``bgk_step(f, tau, force_lattice=F)``,
packages/lattice-boltzmann-d3q19/lattice_boltzmann_d3q19/reference/bgk.py:45-112),
optionally followed each step by the reference's
``apply_bounce_back_y_walls(f, top, bottom)`` (bgk.py:145-216), or with
half-way bounce-back walls (stationary or moving), and optionally a Brinkman
solid fraction s (body force F - s m); u^T = ``macroscopic_velocity(f^T, F - s m)``
(bgk.py:129-142).  Backward: the hand-written adjoint kernel with
checkpoint-every-K recomputation.  Tensors are passed by pointer on the
current HIP stream.
"""

from __future__ import annotations

import ctypes

from . import _device as dev
from ._lib import (
    BOUNDARY,
    D3Q19_F16,
    D3Q19_F32,
    D3Q19_F64,
    FACES,
    FORCE_FIELD,
    STORAGE,
    AdjointOptions,
    Config,
    call,
    load,
)


class DifferentiableLBM:
    """A differentiable D3Q19 forward map (u0, F[, s]) -> u^T on one GPU.

    Parameters
    ----------
    shape, tau, steps : lattice, relaxation time, number of steps T.
    checkpoint_every : K; <= 0 picks the K that needs the least memory.
    storage : "f32" or "f64" -- working precision of forward and adjoint.
    boundary : "periodic", "ywalls" (the reference's y-wall bounce-back with
        ``wall_velocity_top`` / ``wall_velocity_bottom``) or "halfway" (walls on
        ``wall_faces``, moving with ``face_velocity`` {face: (ux, uy, uz)}).
    brinkman : if True the forward map takes a Brinkman solid fraction s.
    checkpoint_storage : None (= storage) or "f16" (f - w checkpoints).
    memory_budget : bytes the adjoint may hold (0 = unlimited); creation
        fails if no checkpoint interval fits.

    Holds the checkpoints of the most recent forward call: one backward per
    forward, in order.  A second backward of the same forward, or the
    backward of a forward that a later one replaced, raises instead of
    returning gradients of the wrong trajectory.
    """

    def __init__(self, shape, tau, steps, checkpoint_every=0, storage="f32", boundary="periodic",
                 wall_faces=(), brinkman=False, checkpoint_storage=None, memory_budget=0, device=0,
                 wall_velocity_top=(0.0, 0.0, 0.0), wall_velocity_bottom=(0.0, 0.0, 0.0), face_velocity=None):
        lib = load()
        self.shape = tuple(int(s) for s in shape)
        self.steps = int(steps)
        self.storage = storage
        self.brinkman = bool(brinkman)
        self.device = int(device)
        cfg = Config()
        lib.d3q19_config_default(ctypes.byref(cfg))
        cfg.nx, cfg.ny, cfg.nz = self.shape
        cfg.tau = float(tau)
        cfg.storage = STORAGE[storage]
        cfg.force_mode = FORCE_FIELD
        try:
            cfg.boundary = BOUNDARY[boundary]
        except KeyError:
            raise ValueError(f"unknown boundary {boundary!r}") from None
        mask = 0
        for face in wall_faces:
            mask |= 1 << FACES[face]
        cfg.wall_faces = mask
        for a in range(3):
            cfg.wall_velocity_top[a] = float(wall_velocity_top[a])
            cfg.wall_velocity_bottom[a] = float(wall_velocity_bottom[a])
        for face, vel in (face_velocity or {}).items():
            for a in range(3):
                cfg.face_velocity[FACES[face]][a] = float(vel[a])
        cfg.device = self.device
        opt = AdjointOptions()
        lib.d3q19_adjoint_options_default(ctypes.byref(opt))
        opt.steps = self.steps
        opt.checkpoint_every = int(checkpoint_every)
        opt.checkpoint_storage = -1 if checkpoint_storage is None else {"f16": D3Q19_F16, **STORAGE}[checkpoint_storage]
        opt.brinkman = 1 if self.brinkman else 0
        opt.memory_budget = int(memory_budget)
        h = ctypes.c_void_p()
        call("d3q19_adjoint_create", ctypes.byref(cfg), ctypes.byref(opt), dev.stream_ptr(self.device),
             ctypes.byref(h))
        self._h = h
        self.tdtype = dev.torch().float64 if storage == "f64" else dev.torch().float32
        self.forward_count = 0  # forwards run; the checkpoints belong to the last one

    @property
    def checkpoint_every(self) -> int:
        return int(load().d3q19_adjoint_checkpoint_every(self._h))

    @property
    def memory_bytes(self) -> int:
        return int(load().d3q19_adjoint_memory_bytes(self._h))

    def close(self):
        if getattr(self, "_h", None):
            load().d3q19_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, t, shape, name):
        if tuple(t.shape) != shape:
            raise ValueError(f"{name} must have shape {shape}, got {tuple(t.shape)}")
        if not t.is_cuda:
            raise ValueError(f"{name} must be a device tensor")

    def _code(self):
        return D3Q19_F64 if self.tdtype == dev.torch().float64 else D3Q19_F32

    def forward_raw(self, u0, force, s=None):
        """u^T for (u0, F[, s]) without autograd bookkeeping (keeps checkpoints)."""
        t = dev.torch()
        v3 = (3,) + self.shape
        self._check(u0, v3, "u0")
        self._check(force, v3, "force")
        u0c = u0.detach().to(self.tdtype).contiguous()
        Fc = force.detach().to(self.tdtype).contiguous()
        sc = None
        if self.brinkman:
            if s is None:
                raise ValueError("this solver takes a Brinkman field s")
            self._check(s, self.shape, "s")
            sc = s.detach().to(self.tdtype).contiguous()
        out = t.empty_like(u0c)
        call("d3q19_adjoint_set_stream", self._h, dev.stream_ptr(self.device))
        self.forward_count += 1
        call("d3q19_adjoint_forward", self._h, dev.ptr(u0c), dev.ptr(Fc), dev.ptr(sc), dev.ptr(out),
             self._code(), 1)
        return out

    def backward_raw(self, grad_u):
        """(dL/du0, dL/dF, dL/ds or None) for dL/du^T of the last forward."""
        t = dev.torch()
        g = grad_u.detach().to(self.tdtype).contiguous()
        gu0 = t.empty_like(g)
        gF = t.empty_like(g)
        gs = t.empty(self.shape, dtype=self.tdtype, device=g.device) if self.brinkman else None
        call("d3q19_adjoint_set_stream", self._h, dev.stream_ptr(self.device))
        call("d3q19_adjoint_backward", self._h, dev.ptr(g), dev.ptr(gu0), dev.ptr(gF), dev.ptr(gs),
             self._code(), 1)
        return gu0, gF, gs

    def __call__(self, u0, force, s=None):
        if self.brinkman:
            return _LBMFunction.apply(u0, force, s, self)
        return _LBMFunction.apply(u0, force, None, self)


def _function_class():
    torch = dev.torch()

    class LBMFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, u0, force, s, solver):
            ctx.solver = solver
            ctx.in_dtypes = (u0.dtype, force.dtype, None if s is None else s.dtype)
            out = solver.forward_raw(u0, force, s)
            ctx.forward_id = solver.forward_count
            return out

        @staticmethod
        def backward(ctx, grad_u):
            if ctx.forward_id != ctx.solver.forward_count:
                raise RuntimeError("DifferentiableLBM holds the checkpoints of its latest forward only: "
                                   "this output's forward was replaced by a later call (one backward per "
                                   "forward, in order)")
            gu0, gF, gs = ctx.solver.backward_raw(grad_u)
            d0, d1, d2 = ctx.in_dtypes
            return gu0.to(d0), gF.to(d1), (None if gs is None else gs.to(d2)), None

    return LBMFunction


class _Lazy:
    _cls = None

    def apply(self, *args):
        if _Lazy._cls is None:
            _Lazy._cls = _function_class()
        return _Lazy._cls.apply(*args)


_LBMFunction = _Lazy()


def velocity_matching_loss(u_final, u_target):
    """mean over cells of |u^T - u_target|^2, accumulated in FP64."""
    t = dev.torch()
    d = (u_final.to(t.float64) - u_target.to(t.float64))
    return (d * d).sum(0).mean()


def outlet_velocity(u_final):
    """Mean x-velocity over the last x-plane (the outlet of a channel along x), FP64."""
    t = dev.torch()
    return u_final[0, -1].to(t.float64).mean()
