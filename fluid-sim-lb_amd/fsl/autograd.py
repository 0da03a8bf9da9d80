"""torch.autograd.Function wrappers of the differentiable projection.

`Projection.apply(ctx, u, params)` projects the face velocities `u` (a
float32 tensor in the box format of fsl.h: shape (nx+2, ny+2, nz+2, 3)) with
the labels already in the context (set by a forward FLIP stage sequence, or
fsl_set_marker) and returns the divergence-free velocities; its backward is
fsl_project_backward, i.e. P^T applied to the incoming gradient with one
additional PCG solve.

A CUDA tensor on the context's device crosses the C ABI in place: the entry
points read and write its storage on the library's stream (after torch's
current stream has been synchronized), so nothing is staged through host
memory. A CPU tensor is copied through host arrays.

In a process that uses both torch's GPU tensors and the library, initialize
torch's GPU runtime (torch.cuda.init()) before the library's first call:
torch's lazy initialization has failed ("No HIP GPUs are available") in a
process where the library had already run rollouts.
"""
from __future__ import annotations

import numpy as np
import torch

from . import Context, SolveParams


def _ones(sim: Context, device) -> torch.Tensor:
    """The all-active mask of the context's box, cached per device."""
    cache = sim.__dict__.setdefault("_ones_mask", {})
    key = str(device)
    if key not in cache:
        cache[key] = torch.ones(sim.box, dtype=torch.uint8, device=device)
    return cache[key]


class Projection(torch.autograd.Function):
    @staticmethod
    def forward(fctx, sim: Context, u: torch.Tensor, params: SolveParams):
        fctx.sim = sim
        fctx.params = params
        fctx.device = u.device
        if u.is_cuda:
            sim.set_vec(0, u, _ones(sim, u.device))
            sim.project_forward(params)
            out, _ = sim.get_vec(0, device=u.device)
            return out
        host = u.detach().to(torch.float32).contiguous().numpy()
        sim.set_vec(0, host, np.ones(sim.box, np.uint8))
        sim.project_forward(params)
        out, _ = sim.get_vec(0)
        return torch.from_numpy(out)

    @staticmethod
    def backward(fctx, grad_out: torch.Tensor):
        if grad_out.is_cuda:
            gin, _ = fctx.sim.project_backward(fctx.params, grad_out)
            return None, gin, None
        g = grad_out.detach().to(torch.float32).contiguous().numpy()
        gin, _ = fctx.sim.project_backward(fctx.params, g)
        return None, torch.from_numpy(gin).to(fctx.device), None


def project(sim: Context, u: torch.Tensor, params: SolveParams) -> torch.Tensor:
    return Projection.apply(sim, u, params)
