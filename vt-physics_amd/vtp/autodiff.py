"""torch.autograd wrappers of the differentiable PBF rollout.

PyTorch is only the thin differentiation layer here: the forward rollout,
its checkpoints and the hand-written adjoint all run in libvtp.so through
the C ABI.  Tensors on the GPU cross the boundary as device pointers
(vtp_write_device / vtp_read_device / vtp_pbf_rollout_bwd_device /
vtp_pbf_density_loss_device: attach-order float[n][3] arrays, no host round
trip); tensors on the CPU go through the host entry points (vtp_write /
vtp_read / vtp_pbf_rollout_bwd / vtp_pbf_density_loss).
"""
from __future__ import annotations

import numpy as np
import torch

from ._abi import Context


def _host(t: torch.Tensor) -> np.ndarray:
    return np.ascontiguousarray(t.detach().to("cpu", torch.float32).numpy())


def _dev(t: torch.Tensor) -> torch.Tensor:
    return t.detach().to(torch.float32).contiguous()


def _ready(device):
    # the library's kernels run on the context's own stream: what torch
    # enqueued on its streams must have finished before they read it
    torch.cuda.synchronize(device)


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, k, model, nsteps, ckpt_every):
        sim = model.sim
        ctx.sim = sim
        ctx.dev = v0.device
        ctx.on_gpu = v0.is_cuda
        sim.set_xsph(float(k.detach()))
        if v0.is_cuda:
            x0 = model.x0_on(v0.device)
            v = _dev(v0)
            _ready(v0.device)
            sim.write_device("pos", x0.data_ptr())
            sim.write_device("vel", v.data_ptr())
            sim.rollout_fwd(nsteps, ckpt_every)
            n = v.shape[0]
            xT = torch.empty((n, 3), dtype=torch.float32, device=v0.device)
            vT = torch.empty((n, 3), dtype=torch.float32, device=v0.device)
            sim.read_device("pos", xT.data_ptr())
            sim.read_device("vel", vT.data_ptr())
            return xT, vT
        sim.write("pos", model.x0)
        sim.write("vel", _host(v0))
        sim.rollout_fwd(nsteps, ckpt_every)
        return torch.from_numpy(sim.read("pos")), torch.from_numpy(sim.read("vel"))

    @staticmethod
    def backward(ctx, gx, gv):
        if ctx.on_gpu:
            gx_d = None if gx is None else _dev(gx)
            gv_d = None if gv is None else _dev(gv)
            n = (gx_d if gx_d is not None else gv_d).shape[0]
            gv0 = torch.empty((n, 3), dtype=torch.float32, device=ctx.dev)
            gk = torch.empty(1, dtype=torch.float32, device=ctx.dev)
            _ready(ctx.dev)
            ctx.sim.rollout_bwd_device(None if gx_d is None else gx_d.data_ptr(),
                                       None if gv_d is None else gv_d.data_ptr(), gv0.data_ptr(), None, gk.data_ptr())
            return gv0, gk.reshape(()), None, None, None
        gx_h = None if gx is None else _host(gx)
        gv_h = None if gv is None else _host(gv)
        gv0, _gx0, gk = ctx.sim.rollout_bwd(gx_h, gv_h)
        return torch.from_numpy(gv0), torch.tensor(gk, dtype=torch.float32), None, None, None


class _DensityLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, sim):
        if x.is_cuda:
            loss = torch.empty(1, dtype=torch.float64, device=x.device)
            g = torch.empty((x.shape[0], 3), dtype=torch.float32, device=x.device)
            _ready(x.device)
            sim.density_loss_device(loss.data_ptr(), g.data_ptr())
            ctx.save_for_backward(g)
            return loss.to(torch.float32).reshape(())
        loss, g = sim.density_loss(with_grad=True)
        ctx.save_for_backward(torch.from_numpy(g))
        return torch.tensor(loss, dtype=torch.float32)

    @staticmethod
    def backward(ctx, go):
        (g,) = ctx.saved_tensors
        return go * g, None


class DifferentiablePBF:
    """PBF rollout x_T, v_T = F(v_0, XSPH_k) with gradients from the adjoint.

    `sim` is an initialised PBF Context; its positions at construction are
    the fixed initial positions of every rollout.  Outputs live where v_0
    lives (the context's GPU, or the CPU).
    """

    def __init__(self, sim: Context, nsteps: int, ckpt_every: int = 10):
        self.sim = sim
        self.nsteps = nsteps
        self.ckpt_every = ckpt_every
        self.x0 = sim.read("pos")
        self.mat = sim.read("mat")
        self._x0_dev = None

    def x0_on(self, device) -> torch.Tensor:
        if self._x0_dev is None or self._x0_dev.device != device:
            self._x0_dev = torch.from_numpy(self.x0).to(device)
        return self._x0_dev

    def rollout(self, v0: torch.Tensor, k: torch.Tensor):
        return _Rollout.apply(v0, k, self, self.nsteps, self.ckpt_every)

    def density_loss(self, x_final: torch.Tensor) -> torch.Tensor:
        """Mean squared density deviation of the fluid at the state of the last rollout."""
        return _DensityLoss.apply(x_final, self.sim)
