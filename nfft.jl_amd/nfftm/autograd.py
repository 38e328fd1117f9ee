"""Differentiation rules of the MI355X NFFT.

The reference differentiates ``plan * x`` and ``nfft(k, x)`` with respect to
the data only: its rrules return ``NoTangent()`` for the plan and the nodes
(AbstractNFFTs/ext/AbstractNFFTsChainRulesCoreExt.jl:14-22, 26-45).  The
pullbacks here return the data cotangent exactly as those rules do
(``adjoint(p) * ybar``) and, in addition, the node cotangent computed on the
device by ``NFFTPlan.grad_nodes`` (analytic derivative of the window).

Cotangent convention for complex values (as ChainRulesCore and PyTorch): for a
real loss L and complex y, ybar = dL/dRe(y) + i dL/dIm(y), so that
dL = Re(conj(ybar) dy).  Then for y = A(k) x:
  xbar = A^H ybar,  kbar[d, j] = Re(conj(ybar_j) dy_j/dk[d, j]).
"""
import numpy as np
import torch

from . import NFFTPlan


def rrule_nfft(k, x, plan=None, **kw):
    """rrule(nfft, k, x): returns (y, pullback) with pullback(ybar) ->
    (kbar, xbar); kbar is a (D, M) float64 array in the axis order of k."""
    if plan is None:
        plan = NFFTPlan(k, tuple(x.shape), **kw)
    y = plan * x
    # the forward pass's transformed grid, kept for the node cotangent (the
    # pullback then runs only the window-derivative gather)
    g = plan.tmpVec.clone()

    def nfft_pullback(ybar):
        kbar = plan.grad_nodes_grid(g, ybar)
        xbar = plan.H * ybar
        return kbar, xbar

    return y, nfft_pullback


class NFFTFunction(torch.autograd.Function):
    """y = A(k) x as a torch autograd function of both k and x.

    ``k``: (D, M) real tensor (any device; its gradient comes back on the same
    device and dtype); ``x``: complex tensor of the plan's image shape.  A plan
    built for these nodes may be passed to avoid rebuilding it.
    """

    @staticmethod
    def forward(ctx, k, x, plan=None, kwargs=None):
        if plan is None:
            plan = NFFTPlan(k.detach().cpu().numpy().astype(np.float64), tuple(x.shape),
                            dtype=torch.float64 if x.dtype == torch.complex128 else torch.float32,
                            **(kwargs or {}))
        ctx.plan = plan
        ctx.k_meta = (k.device, k.dtype)
        ctx.x_dtype = x.dtype
        y = plan * x
        # the transformed grid of this forward pass, for the node cotangent
        ctx.grid = plan.tmpVec.clone() if ctx.needs_input_grad[0] else None
        return y

    @staticmethod
    def backward(ctx, ybar):
        ybar = ybar.contiguous()
        kbar = xbar = None
        if ctx.needs_input_grad[0]:
            dev, dt = ctx.k_meta
            kbar = ctx.plan.grad_nodes_grid(ctx.grid, ybar).to(device=dev, dtype=dt)
        if ctx.needs_input_grad[1]:
            xbar = (ctx.plan.H * ybar).to(ctx.x_dtype)
        return kbar, xbar, None, None


def nfft(k, x, plan=None, **kw):
    """Differentiable nfft(k, x) (AbstractNFFTs/src/derived.jl:134-145)."""
    return NFFTFunction.apply(k, x, plan, kw or None)
