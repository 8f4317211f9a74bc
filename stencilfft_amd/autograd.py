"""Differentiable wrappers over the C ABI (optional torch.autograd binding).

Every operator of the path is linear, so each backward is the operator's
adjoint, computed by the library's adjoint entry points:
  forward FFT          -> zm_forward_adjoint  (the inverse chain, scale 1)
  inverse FFT (s)      -> zm_inverse_adjoint  (s x the forward chain)
  Laplacian            -> itself (symmetric periodic stencil)
  dx4 / dy4 / dz4      -> zm_dx4_adjoint / zm_dy4_adjoint / zm_dz4_adjoint (the
                          transposed stencil, -dx4)
  spectral multiplier  -> zm_spectral_multiply_adjoint (conjugate multiplier;
                          the real kinds are self-adjoint)
For complex tensors torch passes dL/d(conj y) and expects A^H of it back, which
is exactly the adjoint.  No tensors are saved for backward.
"""
from __future__ import annotations

import torch


class _Forward(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, plan):
        ctx.plan = plan
        return plan.forward(x.contiguous())

    @staticmethod
    def backward(ctx, g):
        return ctx.plan.forward_adjoint(g.contiguous()), None


class _Inverse(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, plan, scale):
        ctx.plan, ctx.scale = plan, scale
        return plan.inverse(x.contiguous(), scale)

    @staticmethod
    def backward(ctx, g):
        return ctx.plan.inverse_adjoint(g.contiguous(), ctx.scale), None, None


class _Laplace(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, plan, scale):
        ctx.plan, ctx.scale = plan, scale
        return plan.laplace(x.contiguous(), scale)

    @staticmethod
    def backward(ctx, g):
        return ctx.plan.laplace(g.contiguous(), ctx.scale), None, None


class _Dx4(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, plan, delta):
        ctx.plan, ctx.delta = plan, delta
        return plan.dx4(x.contiguous(), delta)

    @staticmethod
    def backward(ctx, g):
        return ctx.plan.dx4(g.contiguous(), ctx.delta, adjoint=True), None, None


class _Dy4(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, plan, delta):
        ctx.plan, ctx.delta = plan, delta
        return plan.dy4(x.contiguous(), delta)

    @staticmethod
    def backward(ctx, g):
        return ctx.plan.dy4(g.contiguous(), ctx.delta, adjoint=True), None, None


class _Dz4(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, plan, delta):
        ctx.plan, ctx.delta = plan, delta
        return plan.dz4(x.contiguous(), delta)

    @staticmethod
    def backward(ctx, g):
        return ctx.plan.dz4(g.contiguous(), ctx.delta, adjoint=True), None, None


class _Spectral(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, plan, kind, p0, p1, length):
        ctx.args = (plan, kind, p0, p1, length)
        return plan.spectral_multiply(x.contiguous(), kind, p0, p1, length)

    @staticmethod
    def backward(ctx, g):
        plan, kind, p0, p1, length = ctx.args
        return (plan.spectral_multiply(g.contiguous(), kind, p0, p1, length, adjoint=True),
                None, None, None, None, None)


def fft(x, plan):
    """forward_fft: real-space packed layout -> reciprocal layout."""
    return _Forward.apply(x, plan)


def ifft(x, plan, scale=None):
    """inverse_fft with `scale` (default 1/N^d)."""
    if scale is None:
        scale = 1.0 / float(plan.shape.n) ** plan.shape.dims
    return _Inverse.apply(x, plan, float(scale))


def laplace(x, plan, scale=1.0):
    return _Laplace.apply(x, plan, float(scale))


def dx4(x, plan, delta=1.0):
    return _Dx4.apply(x, plan, float(delta))


def dy4(x, plan, delta=1.0):
    return _Dy4.apply(x, plan, float(delta))


def dz4(x, plan, delta=1.0):
    return _Dz4.apply(x, plan, float(delta))


def spectral(x, plan, kind, p0=0.0, p1=0.0, length=1.0):
    return _Spectral.apply(x, plan, kind, float(p0), float(p1), float(length))
