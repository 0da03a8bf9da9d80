"""torch.autograd.Function wrappers of the differentiable projection.

`Projection.apply(ctx, u, params)` projects the face velocities `u` (a
float32 tensor in the box format of fsl.h: shape (nx+2, ny+2, nz+2, 3)) with
the labels already in the context (set by a forward FLIP stage sequence, or
fsl_set_marker) and returns the divergence-free velocities; its backward is
fsl_project_backward, i.e. P^T applied to the incoming gradient with one
additional PCG solve. Tensors cross the C ABI through host memory.
"""
from __future__ import annotations

import numpy as np
import torch

from . import Context, SolveParams


class Projection(torch.autograd.Function):
    @staticmethod
    def forward(fctx, sim: Context, u: torch.Tensor, params: SolveParams):
        host = u.detach().to("cpu", torch.float32).contiguous().numpy()
        sim.set_vec(0, host, np.ones(sim.box, np.uint8))
        sim.project_forward(params)
        out, _ = sim.get_vec(0)
        fctx.sim = sim
        fctx.params = params
        fctx.device = u.device
        return torch.from_numpy(out).to(u.device)

    @staticmethod
    def backward(fctx, grad_out: torch.Tensor):
        g = grad_out.detach().to("cpu", torch.float32).contiguous().numpy()
        gin, _ = fctx.sim.project_backward(fctx.params, g)
        return None, torch.from_numpy(gin).to(fctx.device), None


def project(sim: Context, u: torch.Tensor, params: SolveParams) -> torch.Tensor:
    return Projection.apply(sim, u, params)
