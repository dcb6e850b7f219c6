"""torch.autograd bridge to the adjoint of the Newtonian update.

    W_N = slic_steps(W0, ...)       # torch float64 tensor (ny, nx, 5) or (nx, 5)
    loss(W_N).backward()            # W0.grad from cosmos_adjoint_backward

The forward pass runs n_steps steps of the solve / solve2D loop on the HIP
library (dt from the CFL condition, recorded); the backward pass runs the
hand-written reverse-mode kernels over that trajectory with the recorded dt
schedule held fixed, recomputing each segment between checkpoints.  There is
no CPU fallback: both passes need libcosmos_fv.so and a GPU.
"""
from __future__ import annotations

import numpy as np
import torch

from .solver import VANLEER, Adjoint


class SlicSteps(torch.autograd.Function):
    @staticmethod
    def forward(ctx, W0, order: int, cell_spacing: float, cfl: float, n_steps: int, bias: float, limiter: int,
                checkpoint_every: int, gamma: float, handle_box: list):
        W = W0.detach().to("cpu", torch.float64).contiguous().numpy()
        dim = 2 if W.ndim == 3 else 1
        nx = W.shape[-2]
        ny = W.shape[0] if dim == 2 else 1
        adj = Adjoint(dim, nx, ny, gamma)
        adj.upload(W)
        adj.forward(order, cell_spacing, cfl, n_steps, bias, limiter, checkpoint_every)
        out = adj.download()
        handle_box.append(adj)
        ctx.adj = adj
        ctx.device = W0.device
        return torch.from_numpy(out).to(W0.device)

    @staticmethod
    def backward(ctx, grad_out):
        g = grad_out.detach().to("cpu", torch.float64).contiguous().numpy()
        gi = ctx.adj.backward(g)
        return (torch.from_numpy(gi).to(ctx.device),) + (None,) * 9


def slic_steps(W0: torch.Tensor, cell_spacing: float, cfl: float, n_steps: int, order: int = 2, bias: float = 0.0,
               limiter: int = VANLEER, checkpoint_every: int = 20, gamma: float = 1.4, keep: list | None = None):
    """n_steps solve-loop steps of the Newtonian update as a differentiable
    function of the initial primitive state.  `keep` (a list) receives the
    Adjoint handle, e.g. for replaying the recorded dt schedule."""
    box = keep if keep is not None else []
    return SlicSteps.apply(W0, order, cell_spacing, cfl, n_steps, bias, limiter, checkpoint_every, gamma, box)
