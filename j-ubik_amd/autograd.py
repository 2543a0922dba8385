"""torch.autograd bindings of the HIP operators: the role jax.custom_vjp
plays for the reference's JAX path (the VJP JAX derives for
jubik/convolve.py:66-145 and the likelihood of
jubik/instruments/erosita/erosita_likelihood.py:73-89).

``linpatch_apply(op, x, scale)`` is the PSF operator with its hand-written
adjoint as the backward; ``xray_nll(model, xi_ref, xi_alpha)`` is the fused
Poisson NLL of the X-ray model as a float64 scalar whose backward scales
the gradient the same library call computed with the forward.
"""
import torch

from . import _device as dev


class _Linpatch(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, op, scale):
        ctx.op, ctx.scale = op, scale
        return op.forward(x.detach(), scale)

    @staticmethod
    def backward(ctx, gy):
        return ctx.op.adjoint(gy.contiguous(), ctx.scale), None, None


def linpatch_apply(op, x, scale=1.0):
    """PSF(scale * x) of a ``LinpatchOperator`` (kernel set), differentiable
    in x; the backward is the operator's adjoint."""
    return _Linpatch.apply(dev.to_device(x, op.dtype), op, float(scale))


class _XrayNLL(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xi_ref, xi_alpha, model):
        nll, g_ref, g_alpha = model.nll_and_grad_device(xi_ref.detach(), None if xi_alpha is None
                                                        else xi_alpha.detach())
        ctx.save_for_backward(g_ref.clone(), None if g_alpha is None else g_alpha.clone())
        return nll[0].clone()

    @staticmethod
    def backward(ctx, g):
        g_ref, g_alpha = ctx.saved_tensors
        scale = g.to(g_ref.dtype)
        return scale * g_ref, None if g_alpha is None else scale * g_alpha, None


def xray_nll(model, xi_ref, xi_alpha=None):
    """The fused NLL of an ``XrayModel`` at (xi_ref, xi_alpha) as a float64
    0-d tensor, differentiable in the latent fields (the metric base point
    moves to this position, as for ``nll_and_grad``)."""
    xr = dev.to_device(xi_ref, model.dtype)
    xa = dev.to_device(xi_alpha, model.dtype) if model.sky.has_alpha else None
    return _XrayNLL.apply(xr, xa, model)
