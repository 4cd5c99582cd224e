"""torch.autograd binding of the FMM potentials (a thin layer over the C ABI's fmmh_vjp).

``FmmPotential.apply(points, charges, fmm)`` returns phi (float64, on the operator's GPU)
for points that are both sources and targets; its backward gives dL/dcharges and
dL/dpoints through one more FMM evaluation (the adjoint field of the cotangent), computed
by the library.  The tree of ``fmm`` is the tree of the points it was built for
(``fmm_for``): forward raises when ``points`` differ from them, since the tree and the
leaf coordinates are fixed at build.  ``Potential`` rebuilds the FMM whenever the points
move (an optimisation loop over positions), so each call evaluates at the given points.
Positions are differentiated at fixed tree, which is exact for the smooth far and near
fields.
"""
from __future__ import annotations

import numpy as np
import torch

import fmm_amd


def fmm_for(points, p=6, leaf_points=64, precision="f64", device=0):
    """An Fmm (gradients on) for points given as a (N, 3) tensor or array."""
    pts = points.detach().cpu().double().numpy() if isinstance(points, torch.Tensor) else points
    b = fmm_amd.FmmBuilder(p).gradients(True).max_points_per_leaf(leaf_points).precision(precision).device(device)
    return b.build(pts)


def _host_points(points) -> np.ndarray:
    return points.detach().to("cpu", torch.float64).reshape(-1, 3).numpy()


class FmmPotential(torch.autograd.Function):
    @staticmethod
    def forward(ctx, points, charges, fmm):
        op = fmm.operator
        if not np.array_equal(_host_points(points), fmm.plan.sources):
            raise fmm_amd.FmmError(9, "FmmPotential: the points differ from the points the FMM was built for; "
                                      "build a new one (fmm_for) or use Potential, which rebuilds")
        dev = torch.device("cuda", op.device)
        q = charges.detach().to(dev, torch.float64).contiguous().reshape(-1)
        n = op.ntargets
        phi = torch.empty(n, dtype=torch.float64, device=dev)
        field = torch.empty(n, 3, dtype=torch.float64, device=dev)
        # load_charges orders the operator's stream after torch's current stream, which
        # produced q; read_output_device synchronises the operator before returning
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


class Potential:
    """phi = Potential(p=...)(points, charges): FmmPotential with the FMM rebuilt whenever
    the points differ from the last call's (tree, lists and operator are built for the
    positions given), so gradient steps on the positions see the moved points."""

    def __init__(self, p=6, leaf_points=64, precision="f64", device=0):
        self.settings = dict(p=p, leaf_points=leaf_points, precision=precision, device=device)
        self.fmm = None
        self.builds = 0

    def __call__(self, points, charges):
        if self.fmm is None or not np.array_equal(_host_points(points), self.fmm.plan.sources):
            self.fmm = fmm_for(points, **self.settings)
            self.builds += 1
        return FmmPotential.apply(points, charges, self.fmm)
