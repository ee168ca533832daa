"""Differentiable batched rollout: torch.autograd over the C ABI's hand-written
adjoint kernels (include/sph2s.h, sph_rollout_*).

``Rollout`` owns the device state of one batched rollout (scene table,
checkpoints, segment records).  ``rollout(r, pos0, vel0, stiffness,
viscosity)`` runs it forward and returns the final positions as a torch tensor
on the GPU; ``backward`` through it yields gradients for the initial positions,
the initial velocities and every scene's stiffness and viscosity.  Topology
(which particles interact) is frozen within each step, as the forward sees it.
Tensors are (N, 3) float32 on the GPU in global particle order (scene-major);
stiffness and viscosity are (n_scenes,) float32 tensors on any device.  The
work is ordered on torch's current stream.  There is no CPU path: without the
built library or a GPU every call raises.
"""
from __future__ import annotations

from ctypes import byref, c_size_t, c_uint32, c_void_p

import numpy as np
import torch

from . import _capi as C


def _ptr(t):
    return None if t is None else c_void_p(t.data_ptr())


def _stream():
    return c_void_p(torch.cuda.current_stream().cuda_stream)


class Rollout:
    """n_steps steps of dt over a batch of scenes, checkpointed every
    ``checkpoint_every`` steps (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:1082-1198 is
    the step being differentiated)."""

    def __init__(self, configurations, counts, dt: float, n_steps: int,
                 checkpoint_every: int = 20):
        cfgs = list(configurations)
        self.counts = np.ascontiguousarray(counts, np.uint32)
        if len(cfgs) != self.counts.size:
            raise ValueError("one configuration per scene")
        arr = (C.SphConfig * len(cfgs))(*cfgs)
        self._h = c_void_p()
        C.check(C.lib().sph_rollout_create(arr, len(cfgs), self.counts.ctypes.data_as(c_void_p),
                                           float(dt), int(n_steps), int(checkpoint_every),
                                           byref(self._h)))
        n, ns, nb = c_uint32(), c_uint32(), c_size_t()
        C.check(C.lib().sph_rollout_info(self._h, byref(n), byref(ns), byref(nb)))
        self.n, self.n_scenes, self.device_bytes = int(n.value), int(ns.value), int(nb.value)
        self.dt, self.n_steps, self.checkpoint_every = float(dt), int(n_steps), int(checkpoint_every)
        self.configurations = cfgs

    def _check_state(self, t, name):
        if t is None:
            return None
        if not (t.is_cuda and t.dtype == torch.float32 and tuple(t.shape) == (self.n, 3)):
            raise ValueError(f"{name} must be a ({self.n}, 3) float32 CUDA tensor")
        return t.contiguous()

    @staticmethod
    def _params(t, n_scenes):
        if t is None:
            return None
        a = np.ascontiguousarray(t.detach().to("cpu", torch.float32).numpy().reshape(-1))
        if a.size != n_scenes:
            raise ValueError("one parameter value per scene")
        return a

    def forward(self, pos0, vel0=None, stiffness=None, viscosity=None) -> torch.Tensor:
        pos0 = self._check_state(pos0, "pos0")
        vel0 = self._check_state(vel0, "vel0")
        k = self._params(stiffness, self.n_scenes)
        mu = self._params(viscosity, self.n_scenes)
        out = torch.empty_like(pos0)
        C.check(C.lib().sph_rollout_forward(
            self._h, _ptr(pos0), _ptr(vel0), None if k is None else k.ctypes.data_as(c_void_p),
            None if mu is None else mu.ctypes.data_as(c_void_p), _ptr(out), _stream()))
        return out

    def backward(self, grad_final):
        g = self._check_state(grad_final, "grad_final")
        gp = torch.empty_like(g)
        gv = torch.empty_like(g)
        gk = np.zeros(self.n_scenes, np.float64)
        gm = np.zeros(self.n_scenes, np.float64)
        C.check(C.lib().sph_rollout_backward(self._h, _ptr(g), _ptr(gp), _ptr(gv),
                                             gk.ctypes.data_as(c_void_p),
                                             gm.ctypes.data_as(c_void_p), _stream()))
        return gp, gv, gk, gm

    def synchronize(self) -> None:
        C.check(C.lib().sph_rollout_synchronize(self._h))

    def device_errors(self) -> int:
        bits = c_uint32()
        C.check(C.lib().sph_rollout_device_errors(self._h, byref(bits)))
        return int(bits.value)

    def close(self) -> None:
        if self._h:
            C.lib().sph_rollout_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class RolloutFunction(torch.autograd.Function):
    """x_T = rollout(x0, v0; stiffness, viscosity).  One backward per forward:
    the rollout's checkpoints belong to the latest forward call."""

    @staticmethod
    def forward(ctx, r: Rollout, pos0, vel0, stiffness, viscosity):
        ctx.r = r
        ctx.param_meta = [(t.dtype, t.device) for t in (stiffness, viscosity)]
        return r.forward(pos0, vel0, stiffness, viscosity)

    @staticmethod
    def backward(ctx, grad_out):
        gp, gv, gk, gm = ctx.r.backward(grad_out)
        (kd, kdev), (md, mdev) = ctx.param_meta
        return (None, gp, gv, torch.from_numpy(gk).to(kdev, kd),
                torch.from_numpy(gm).to(mdev, md))


def rollout(r: Rollout, pos0, vel0, stiffness, viscosity) -> torch.Tensor:
    """Differentiable final positions of the batched rollout ``r``."""
    return RolloutFunction.apply(r, pos0, vel0, stiffness, viscosity)
