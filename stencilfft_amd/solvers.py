"""Differentiable spectral / stencil solves built from the path's operators.

poisson_solve    u = IFFT(FFT(f) / (-|k|^2)), mean mode removed (multiplier in FP64)
poisson_loss     ||Laplacian_5pt(u) + dx4(u) - g||^2 with h = delta = 1/N
                 (poisson_residual returns the residual itself).  form="spatial"
                 runs the stencil kernels on u; form="spectral" applies their exact
                 symbols fused with the Poisson multiplier in reciprocal space.
                 Both are the same linear map; the spatial form in FP32 is
                 ill-conditioned at large N: the N^2-scaled Laplacian lifts the
                 FP32 rounding of the smooth u (white, ~1e-7 |u|) by up to 8 N^2,
                 so its FP32 residual carries ~1e-2 relative error at N = 4096,
                 while the spectral form keeps FP32 at ~1e-7.
heat_step        u <- IFFT(exp(-nu |k|^2 dt) FFT(u)) + dt * Laplacian_7pt(u)
rollout          n steps of any step function with checkpoints every k steps:
                 the forward keeps only the checkpointed states, the backward
                 recomputes each segment from its checkpoint and runs the
                 adjoint kernels through it (peak memory ~ n/k + k states).
adjoint_heat     the same gradient as a hand-written adjoint loop
                 (u_T^H-misfit pulled back by the adjoint step operator).
"""
from __future__ import annotations

import torch

from . import autograd as ag


def poisson_solve(f, plan, length=1.0):
    y = ag.fft(f, plan)
    y = ag.spectral(y, plan, "poisson", length=length)
    return ag.ifft(y, plan)


def poisson_residual(f, g, plan, length=1.0, form="spatial"):
    """Laplacian_5pt(u) + dx4(u) - g with u = poisson_solve(f), h = delta = 1/N."""
    n = plan.shape.n
    if form == "spectral":
        y = ag.spectral(ag.fft(f, plan), plan, "poisson_stencil", float(n * n), 1.0 / n, length)
        return ag.ifft(y, plan) - g
    if form != "spatial":
        raise ValueError(f"unknown form {form!r}")
    u = poisson_solve(f, plan, length)
    return ag.laplace(u, plan, scale=float(n * n)) + ag.dx4(u, plan, delta=1.0 / n) - g


def sq_norm(r):
    """||r||^2 with a well-defined complex gradient (2 r)."""
    return torch.view_as_real(r).pow(2).sum()


def poisson_loss(f, g, plan, length=1.0, form="spatial"):
    return sq_norm(poisson_residual(f, g, plan, length, form))


def heat_step(u, plan, nu, dt, lap_scale=1.0, length=1.0):
    y = ag.spectral(ag.fft(u, plan), plan, "heat", nu, dt, length)
    return ag.ifft(y, plan) + dt * ag.laplace(u, plan, scale=lap_scale)


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, step, nsteps, every):
        ctx.step, ctx.nsteps, ctx.every = step, nsteps, every
        checkpoints = []
        u = u0
        with torch.no_grad():
            for t in range(nsteps):
                if t % every == 0:
                    checkpoints.append(u)
                u = step(u)
        ctx.checkpoints = checkpoints
        return u

    @staticmethod
    def backward(ctx, g):
        step, nsteps, every = ctx.step, ctx.nsteps, ctx.every
        for seg in reversed(range(len(ctx.checkpoints))):
            t0 = seg * every
            t1 = min(t0 + every, nsteps)
            with torch.enable_grad():
                u = ctx.checkpoints[seg].detach().requires_grad_(True)
                v = u
                for _ in range(t0, t1):
                    v = step(v)
                (g,) = torch.autograd.grad(v, u, grad_outputs=g)
        ctx.checkpoints = None
        return g, None, None, None


def rollout(u0, step, nsteps, every=8):
    """nsteps applications of `step` with checkpoints every `every` steps."""
    return _Rollout.apply(u0, step, int(nsteps), int(every))


def adjoint_heat(g, plan, nsteps, nu, dt, lap_scale=1.0, length=1.0):
    """Pull a final-state gradient back through nsteps heat steps with the
    adjoint kernels directly: step^H = F^H M IFFT^H + dt Lap (M real, Lap symmetric)."""
    n_d = float(plan.shape.n) ** plan.shape.dims
    for _ in range(nsteps):
        a = plan.inverse_adjoint(g.contiguous(), 1.0 / n_d)
        a = plan.spectral_multiply(a, "heat", nu, dt, length)
        a = plan.forward_adjoint(a)
        g = a + dt * plan.laplace(g.contiguous(), lap_scale)
    return g
