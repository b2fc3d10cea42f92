"""Differentiable PM: forward leapfrog with per-step checkpoints and the hand-written
adjoint of ``lf2mi_diffpm_*`` (include/lf2mi.h), plus a ``torch.autograd.Function``
harness around it.

No counterpart in the reference: it differentiates the reference's PM cycle
(scalarProjectionCIC_project particles/projections.hpp:76-300, the Poisson solve of
examples/poissonSolver.cpp:99-111, updateVel/moveParticles Particles.hpp:825-828,
:2028-2032).  Box length 1, periodic; positions in box units, particle order kept.

    model = DiffPM(N=128, np=128**3, nsteps=10, mass=1/128**3, dt=0.02, dtype=F32)
    loss = pm_density_loss(pos0, vel0, target, model)   # pos0.requires_grad_()
    loss.backward()                                      # pos0.grad = dL/dpos0
"""
from __future__ import annotations

import ctypes

from . import _lib
from ._lib import F32, F64, call


def _torch():
    import torch
    return torch


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class DiffPM:
    """nsteps of the kick-drift PM leapfrog on an N^3 mesh with np particles of mass
    `mass` (density mass*N^3 per particle), time step dt, computed in `dtype`."""

    def __init__(self, N: int, np: int, nsteps: int, mass: float, dt: float, dtype=F32):
        self.N, self.np, self.nsteps = int(N), int(np), int(nsteps)
        self.mass, self.dt = float(mass), float(dt)
        self.code = F32 if dtype in (F32, "f32") or str(dtype) == "torch.float32" else F64
        h = ctypes.c_void_p()
        call("lf2mi_diffpm_create", self.N, self.np, self.nsteps, self.mass, self.dt, self.code, ctypes.byref(h))
        self.h = h
        self._io = None
        # counts the forwards: the device holds the checkpoints of the last one only
        self.generation = 0

    def __del__(self):
        try:
            if getattr(self, "h", None) is not None and _lib._lib is not None:
                _lib._lib.lf2mi_diffpm_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def bytes(self) -> int:
        n = ctypes.c_int64(0)
        call("lf2mi_diffpm_bytes", self.h, ctypes.byref(n))
        return n.value

    @staticmethod
    def _io_code(t):
        torch = _torch()
        if t.dtype == torch.float32:
            return F32
        if t.dtype == torch.float64:
            return F64
        raise TypeError(f"positions must be float32 or float64, not {t.dtype}")

    def _check(self, t, n, what):
        if t is None:
            return None
        if not t.is_cuda or not t.is_contiguous() or t.numel() != n:
            raise ValueError(f"{what}: expected a contiguous device tensor of {n} values")
        return t

    def forward(self, pos0, vel0=None, target=None, want_rho: bool = False):
        """loss (float) and, with want_rho, the final density (N,N,N) indexed [z,y,x]"""
        torch = _torch()
        io = self._io_code(pos0)
        self._check(pos0, 3 * self.np, "pos0")
        for t, n, w in ((vel0, 3 * self.np, "vel0"), (target, self.N ** 3, "target")):
            self._check(t, n, w)
            if t is not None and t.dtype != pos0.dtype:
                raise TypeError(f"{w} must have the dtype of pos0")
        rho = torch.empty((self.N,) * 3, dtype=pos0.dtype, device=pos0.device) if want_rho else None
        loss = ctypes.c_double(0.0)
        call("lf2mi_diffpm_forward", self.h, _ptr(pos0), _ptr(vel0), _ptr(target), io, _ptr(rho),
             ctypes.byref(loss), _stream())
        self._io = pos0.dtype
        self.generation += 1
        return (loss.value, rho) if want_rho else loss.value

    def backward(self, want_vel: bool = True, generation: int | None = None):
        """dL/dpos0 and dL/dvel0 of the last forward, shaped (np, 3); with `generation`,
        raise unless the last forward is that one"""
        torch = _torch()
        if self._io is None:
            raise RuntimeError("backward before forward")
        if generation is not None and generation != self.generation:
            raise RuntimeError(f"the model's checkpoints belong to forward {self.generation}, not to forward "
                               f"{generation}: run each forward's backward before the next forward on the same "
                               "DiffPM, or use one DiffPM per evaluation")
        gp = torch.empty((self.np, 3), dtype=self._io, device="cuda")
        gv = torch.empty((self.np, 3), dtype=self._io, device="cuda") if want_vel else None
        call("lf2mi_diffpm_backward", self.h, _ptr(gp), _ptr(gv), _stream())
        return gp, gv

    def state(self, step: int, dtype=None):
        """(pos, vel) in box units at the start of `step` of the last forward"""
        torch = _torch()
        dt = dtype or self._io or torch.float64
        pos = torch.empty((self.np, 3), dtype=dt, device="cuda")
        vel = torch.empty((self.np, 3), dtype=dt, device="cuda")
        call("lf2mi_diffpm_state", self.h, int(step), _ptr(pos), _ptr(vel), self._io_code(pos), _stream())
        return pos, vel


def _autograd_function():
    torch = _torch()

    class PMDensityLoss(torch.autograd.Function):
        """L = sum_cells (rho_final - target)^2 through the GPU leapfrog; the backward
        runs the hand-written adjoint kernels (not torch autograd).  The model keeps the
        checkpoints of its last forward only, so a backward whose forward was followed by
        another forward of the same model raises instead of differentiating the wrong
        trajectory."""

        @staticmethod
        def forward(ctx, pos0, vel0, target, model):
            loss, rho = model.forward(pos0.detach().contiguous(),
                                      None if vel0 is None else vel0.detach().contiguous(),
                                      None if target is None else target.detach().contiguous(), want_rho=True)
            ctx.model = model
            ctx.generation = model.generation
            ctx.pos_shape = pos0.shape
            ctx.has_vel = vel0 is not None
            ctx.has_target = target is not None
            ctx.save_for_backward(rho, target if target is not None else torch.zeros_like(rho))
            return torch.tensor(loss, dtype=torch.float64, device=pos0.device)

        @staticmethod
        def backward(ctx, g):
            rho, target = ctx.saved_tensors
            gp, gv = ctx.model.backward(want_vel=ctx.has_vel, generation=ctx.generation)
            s = g.to(gp.dtype)
            gt = (-2.0 * (rho - target.reshape(rho.shape)) * s).reshape(target.shape) if ctx.has_target else None
            return (gp * s).reshape(ctx.pos_shape), (gv * s).reshape(ctx.pos_shape) if gv is not None else None, gt, None

    return PMDensityLoss


_FN = None


def pm_density_loss(pos0, vel0, target, model: DiffPM):
    """differentiable in pos0, vel0 (may be None) and target (may be None)"""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(pos0, vel0, target, model)
