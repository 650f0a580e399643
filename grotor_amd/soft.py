"""torch.autograd.Function over the soft (log-sum-exp) L-shape DP.

The forward and the hand-written adjoint run in libgrotor_amd.so
(gr_soft_forward / gr_soft_backward); torch only carries the three
cost-model parameters (overflow slope, zero-capacity slope, via weight) and
the scalar result.  The adjoint reuses the tables the forward checkpointed on
the device, so a backward must follow its own forward on the same context.
"""
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


def soft_lshape_cost(context, params, begin, count, tau=1.0, precision=32):
    """Differentiable soft L-shape cost of a batch of stage slots."""
    return SoftLshapeCost.apply(params, context, begin, count, tau, precision)
