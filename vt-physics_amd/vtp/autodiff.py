"""torch.autograd wrappers of the differentiable PBF rollout.

PyTorch is only the thin differentiation layer here: the forward rollout,
its checkpoints and the hand-written adjoint all run in libvtp.so through
the C ABI (vtp_pbf_rollout_fwd / vtp_pbf_rollout_bwd / vtp_pbf_density_loss).
Tensors cross the boundary as host float32 arrays in attach order.
"""
from __future__ import annotations

import numpy as np
import torch

from ._abi import Context


def _host(t: torch.Tensor) -> np.ndarray:
    return np.ascontiguousarray(t.detach().to("cpu", torch.float32).numpy())


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, k, sim, x0, nsteps, ckpt_every):
        sim.write("pos", x0)
        sim.write("vel", _host(v0))
        sim.set_xsph(float(k.detach().cpu()))
        sim.rollout_fwd(nsteps, ckpt_every)
        ctx.sim = sim
        ctx.dev = v0.device
        x = torch.from_numpy(sim.read("pos")).to(v0.device)
        v = torch.from_numpy(sim.read("vel")).to(v0.device)
        return x, v

    @staticmethod
    def backward(ctx, gx, gv):
        gx_h = None if gx is None else _host(gx)
        gv_h = None if gv is None else _host(gv)
        gv0, _gx0, gk = ctx.sim.rollout_bwd(gx_h, gv_h)
        return (torch.from_numpy(gv0).to(ctx.dev), torch.tensor(gk, dtype=torch.float32, device=ctx.dev),
                None, None, None, None)


class _DensityLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, sim):
        loss, g = sim.density_loss(with_grad=True)
        ctx.save_for_backward(torch.from_numpy(g).to(x.device))
        return torch.tensor(loss, dtype=torch.float32, device=x.device)

    @staticmethod
    def backward(ctx, go):
        (g,) = ctx.saved_tensors
        return go * g, None


class DifferentiablePBF:
    """PBF rollout x_T, v_T = F(v_0, XSPH_k) with gradients from the adjoint.

    `sim` is an initialised PBF Context; its positions at construction are
    the fixed initial positions of every rollout.
    """

    def __init__(self, sim: Context, nsteps: int, ckpt_every: int = 10):
        self.sim = sim
        self.nsteps = nsteps
        self.ckpt_every = ckpt_every
        self.x0 = sim.read("pos")
        self.mat = sim.read("mat")

    def rollout(self, v0: torch.Tensor, k: torch.Tensor):
        return _Rollout.apply(v0, k, self.sim, self.x0, self.nsteps, self.ckpt_every)

    def density_loss(self, x_final: torch.Tensor) -> torch.Tensor:
        """Mean squared density deviation of the fluid at the state of the last rollout."""
        return _DensityLoss.apply(x_final, self.sim)
