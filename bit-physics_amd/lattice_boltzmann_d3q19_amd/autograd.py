"""torch.autograd.Function over the engine's adjoint (C ABI d3q19_adjoint_*).

    solver = DifferentiableLBM((128, 128, 128), tau=0.6, steps=2000, checkpoint_every=50)
    u_T = solver(u0, F)                 # u0, F: (3, Nx, Ny, Nz) device tensors
    loss = ((u_T - u_target) ** 2).sum(0).mean()
    loss.backward()                     # u0.grad, F.grad via the adjoint kernels

Forward: f^0 = feq(1, u0); ``steps`` periodic BGK+Guo steps with the per-cell
force F (the reference's ``bgk_step(f, tau, force_lattice=F)``,
packages/lattice-boltzmann-d3q19/lattice_boltzmann_d3q19/reference/bgk.py:45-112);
u^T = ``macroscopic_velocity(f^T, F)`` (bgk.py:129-142).  Backward: the
hand-written adjoint kernel with checkpoint-every-K recomputation.
Tensors are passed by pointer on the current HIP stream.
"""

from __future__ import annotations

import ctypes

from . import _device as dev
from ._lib import D3Q19_F32, D3Q19_F64, FORCE_FIELD, STORAGE, Config, call, load


class DifferentiableLBM:
    """A differentiable D3Q19 forward map u0, F -> u^T on one GPU.

    Holds the checkpoints of the most recent forward call: one backward per
    forward, in order.
    """

    def __init__(self, shape, tau, steps, checkpoint_every=0, storage="f32", device=0):
        lib = load()
        self.shape = tuple(int(s) for s in shape)
        self.steps = int(steps)
        self.storage = storage
        self.device = int(device)
        cfg = Config()
        lib.d3q19_config_default(ctypes.byref(cfg))
        cfg.nx, cfg.ny, cfg.nz = self.shape
        cfg.tau = float(tau)
        cfg.storage = STORAGE[storage]
        cfg.force_mode = FORCE_FIELD
        cfg.device = self.device
        h = ctypes.c_void_p()
        call("d3q19_adjoint_create", ctypes.byref(cfg), self.steps, int(checkpoint_every),
             dev.stream_ptr(self.device), ctypes.byref(h))
        self._h = h
        self.tdtype = dev.torch().float64 if storage == "f64" else dev.torch().float32

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

    def _check(self, t, name):
        if tuple(t.shape) != (3,) + self.shape:
            raise ValueError(f"{name} must have shape {(3,) + self.shape}, got {tuple(t.shape)}")
        if not t.is_cuda:
            raise ValueError(f"{name} must be a device tensor")

    def forward_raw(self, u0, force):
        """u^T for (u0, F) without autograd bookkeeping (keeps checkpoints)."""
        t = dev.torch()
        self._check(u0, "u0")
        self._check(force, "force")
        u0c = u0.detach().to(self.tdtype).contiguous()
        Fc = force.detach().to(self.tdtype).contiguous()
        out = t.empty_like(u0c)
        code = D3Q19_F64 if self.tdtype == t.float64 else D3Q19_F32
        call("d3q19_adjoint_set_stream", self._h, dev.stream_ptr(self.device))
        call("d3q19_adjoint_forward", self._h, dev.ptr(u0c), dev.ptr(Fc), dev.ptr(out), code, 1)
        return out

    def backward_raw(self, grad_u):
        t = dev.torch()
        g = grad_u.detach().to(self.tdtype).contiguous()
        gu0 = t.empty_like(g)
        gF = t.empty_like(g)
        code = D3Q19_F64 if self.tdtype == t.float64 else D3Q19_F32
        call("d3q19_adjoint_set_stream", self._h, dev.stream_ptr(self.device))
        call("d3q19_adjoint_backward", self._h, dev.ptr(g), dev.ptr(gu0), dev.ptr(gF), code, 1)
        return gu0, gF

    def __call__(self, u0, force):
        return _LBMFunction.apply(u0, force, self)


def _function_class():
    torch = dev.torch()

    class LBMFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, u0, force, solver):
            ctx.solver = solver
            ctx.in_dtypes = (u0.dtype, force.dtype)
            return solver.forward_raw(u0, force)

        @staticmethod
        def backward(ctx, grad_u):
            gu0, gF = ctx.solver.backward_raw(grad_u)
            return gu0.to(ctx.in_dtypes[0]), gF.to(ctx.in_dtypes[1]), None

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
