"""torch.autograd binding of the FMM potentials (a thin layer over the C ABI's fmmh_vjp).

``FmmPotential.apply(points, charges, fmm)`` returns phi (float64, on cuda:0) for points
that are both sources and targets; its backward gives dL/dcharges and dL/dpoints through
one more FMM evaluation (the adjoint field of the cotangent), computed by the library.
The tree of ``fmm`` must be the tree of ``points`` (build it with ``fmm_for``); positions
are differentiated at fixed tree, which is exact for the smooth far and near fields.
"""
from __future__ import annotations

import torch

import fmm_amd


def fmm_for(points, p=6, leaf_points=64, precision="f64"):
    """An Fmm (gradients on) for points given as a (N, 3) tensor or array."""
    pts = points.detach().cpu().double().numpy() if isinstance(points, torch.Tensor) else points
    return fmm_amd.FmmBuilder(p).gradients(True).max_points_per_leaf(leaf_points).precision(precision).build(pts)


class FmmPotential(torch.autograd.Function):
    @staticmethod
    def forward(ctx, points, charges, fmm):
        dev = torch.device("cuda", 0)
        q = charges.detach().to(dev, torch.float64).contiguous()
        n = q.numel()
        phi = torch.empty(n, dtype=torch.float64, device=dev)
        field = torch.empty(n, 3, dtype=torch.float64, device=dev)
        op = fmm.operator
        op.load_charges(q)
        op.reset()
        op.evaluate()
        op.read_output_device(phi, field)
        ctx.fmm = fmm
        ctx.points_device = points.device
        ctx.charges_device = charges.device
        ctx.save_for_backward(q, field)
        return phi

    @staticmethod
    def backward(ctx, gphi):
        q, field = ctx.saved_tensors
        g = gphi.detach().to(q.device, torch.float64).contiguous()
        gq = torch.empty_like(q)
        gp = torch.empty_like(field)
        ctx.fmm.operator.vjp_device(q, g, field, gq, gp)
        return gp.to(ctx.points_device), gq.to(ctx.charges_device), None
