"""Seeded synthetic fields of the benchmark configurations, generated on the GPU
and delivered in the real-space packed layout (the analogue of init_host,
zmodel/utils.cu:9-81, for random data).

mode field: sum of `nmodes` plane waves exp(2 pi i k.x / N) with integer
wavevectors |k| <= kmax drawn uniformly from the ball, amplitudes
N(0,1) / (1 + |k|^2) and uniform phases, plus `noise` * U[-1, 1) in re and im.
Wavevectors, amplitudes and phases come from numpy's PCG64 default_rng(seed);
the noise from torch's CUDA generator seeded with `seed`.  Phases are reduced
exactly ((k x) mod N) before the exponential, and the sum is evaluated in FP64
as rank-`nmodes` complex matrix products.
"""
from __future__ import annotations

import numpy as np
import torch


def mode_parameters(dims: int, nmodes: int, kmax: int, seed: int):
    rng = np.random.default_rng(seed)
    ks, coef = [], []
    while len(ks) < nmodes:
        k = rng.integers(-kmax, kmax + 1, size=dims)
        k2 = float(np.sum(k.astype(np.float64) ** 2))
        if k2 > kmax * kmax:
            continue
        ks.append(k)
        coef.append(rng.normal() / (1.0 + k2) * np.exp(1j * rng.uniform(0, 2 * np.pi)))
    return np.array(ks), np.array(coef), rng


def _waves(n, kcol):
    pos = np.arange(n)
    return np.exp(2j * np.pi * ((kcol[None, :] * pos[:, None]) % n) / n)  # (n, nmodes)


def mode_field_natural(n: int, dims: int, nmodes: int, kmax: int, seed: int, noise: float = 0.0):
    """complex128 CUDA tensor of N^dims samples in natural row-major order
    (axis 0 = rows gr in 2D / depth gd in 3D, last axis = columns gc)."""
    ks, coef, _ = mode_parameters(dims, nmodes, kmax, seed)
    if dims == 2:
        U = torch.from_numpy(_waves(n, ks[:, 0]) * coef[None, :]).cuda()
        V = torch.from_numpy(np.ascontiguousarray(_waves(n, ks[:, 1]).T)).cuda()
        f = U @ V
    else:
        Z = torch.from_numpy(_waves(n, ks[:, 0]) * coef[None, :]).cuda()  # (n, m)
        Y = torch.from_numpy(_waves(n, ks[:, 1])).cuda()                   # (n, m)
        X = torch.from_numpy(np.ascontiguousarray(_waves(n, ks[:, 2]).T)).cuda()  # (m, n)
        f = torch.empty((n, n, n), dtype=torch.complex128, device="cuda")
        for z in range(n):
            f[z] = (Y * Z[z][None, :]) @ X
    if noise:
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        shape = (n,) * dims
        re = torch.rand(shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
        im = torch.rand(shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
        f += noise * torch.complex(re, im)
    return f.reshape(-1).contiguous()


def mode_fields(plan, nmodes: int, kmax: int, seed: int, noise: float = 0.0):
    """`plan.batch` mode fields (seeds seed, seed+1, ...) in the packed layout,
    complex128 for an FP64 plan and the correctly rounded complex64 otherwise."""
    n, dims = plan.shape.n, plan.shape.dims
    nat = torch.empty(plan.numel(), dtype=plan.complex_dtype, device="cuda")
    field = plan.dims.field
    for i in range(plan.batch):
        nat[i * field:(i + 1) * field] = mode_field_natural(n, dims, nmodes, kmax, seed + i, noise)
    return plan.natural_to_packed(nat)


def uniform_fields(plan, seed: int):
    """re, im i.i.d. U[-1, 1) (torch CUDA generator), packed layout of `plan`."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    re = torch.rand(plan.numel(), generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    im = torch.rand(plan.numel(), generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    return torch.complex(re, im).to(plan.complex_dtype).contiguous()
