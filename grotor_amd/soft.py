"""torch.autograd.Functions over the soft (log-sum-exp) DPs.

The forward and the hand-written adjoint run in libgrotor_amd.so
(gr_soft_forward / gr_soft_forward_costs and gr_soft_backward); torch only
carries the inputs and the scalar result.  SoftLshapeCost takes the three
cost-model parameters (overflow slope, zero-capacity slope, via weight);
SoftDagCost takes the per-cell wire and via cost tensors themselves, whose
gradients are the expected wire-edge and via-landing use.  The adjoint
reuses the tables the forward checkpointed on the device, so a backward must
follow its own forward on the same context.
"""
import numpy as np
import torch


class SoftLshapeCost(torch.autograd.Function):
    """total = sum over the stage slots [begin, begin + count) of the soft DP
    value, as a function of params = (slope, zero_slope, via_weight)."""

    @staticmethod
    def forward(fctx, params, context, begin, count, tau, precision):
        slope, zero_slope, via_weight = params.detach().to("cpu", torch.float64).tolist()
        total, _, _ = context.soft_forward(begin, count, slope, zero_slope, via_weight, tau, precision)
        fctx.context = context
        fctx.generation = context.soft_generation
        return torch.tensor(total, dtype=torch.float64, device=params.device)

    @staticmethod
    def backward(fctx, grad_total):
        context = fctx.context
        if context.soft_generation != fctx.generation:
            raise RuntimeError("the soft DP tables of this forward were overwritten by a later forward")
        grad_params, _ = context.soft_backward()
        g = torch.as_tensor(grad_params, dtype=torch.float64, device=grad_total.device) * grad_total
        return g, None, None, None, None, None


class SoftDagCost(torch.autograd.Function):
    """total = sum over the stage slots [begin, begin + count) of the soft DP
    value of the context's edge kind (context.set_soft_kind), as a function
    of the per-cell wire costs and via costs (float64 tensors of L * X * Y
    cells, IDX order)."""

    @staticmethod
    def forward(fctx, wcost, vcost, context, begin, count, tau, precision):
        w = wcost.detach().to("cpu", torch.float64).numpy()
        v = vcost.detach().to("cpu", torch.float64).numpy()
        total, _, _ = context.soft_forward_costs(begin, count, w, v, tau, precision)
        fctx.context = context
        fctx.generation = context.soft_generation
        fctx.devices = (wcost.device, vcost.device)
        return torch.tensor(total, dtype=torch.float64, device=wcost.device)

    @staticmethod
    def backward(fctx, grad_total):
        context = fctx.context
        if context.soft_generation != fctx.generation:
            raise RuntimeError("the soft DP tables of this forward were overwritten by a later forward")
        context.soft_backward()
        gw, gv = context.soft_edge_grads()
        scale = grad_total.detach().to("cpu", torch.float64)
        gw = torch.from_numpy(np.ascontiguousarray(gw)) * scale
        gv = torch.from_numpy(np.ascontiguousarray(gv)) * scale
        return gw.to(fctx.devices[0]), gv.to(fctx.devices[1]), None, None, None, None, None


def soft_lshape_cost(context, params, begin, count, tau=1.0, precision=32):
    """Differentiable soft L-shape cost of a batch of stage slots."""
    return SoftLshapeCost.apply(params, context, begin, count, tau, precision)


def soft_dag_cost(context, wcost, vcost, begin, count, tau=1.0, precision=64, kind=None):
    """Differentiable soft DP cost of a batch of stage slots over given
    wire and via cost tensors; kind (0 L-shape, 1 Z-shape) sets the
    context's edge kind first when given."""
    if kind is not None:
        context.set_soft_kind(kind)
    return SoftDagCost.apply(wcost, vcost, context, begin, count, tau, precision)
