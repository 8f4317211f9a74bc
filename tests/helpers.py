"""Shared test helpers: golden-fixture parsing, natural-layout reference
stencils written from their textbook definitions, and seeded field generators."""
from __future__ import annotations

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# Relative max-error bars (largest |error| over largest |reference|) of the two
# FP32 paths against FP64; both also keep the 1e-6 relative-L2 bar.
FP32_REL_MAX = 1e-5
MFMA_REL_MAX = 1e-5


def load_images():
    """Parse tests/golden/images_n16_p2_b2.txt -> {step: {rank_name: (8, 8) int array}}."""
    steps: dict = {}
    cur = None
    with open(os.path.join(GOLDEN, "images_n16_p2_b2.txt")) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            if line.startswith("step"):
                st, rk = line.split()
                cur = steps.setdefault(st, {}).setdefault(rk, [])
                continue
            cur.append([int(x) for x in line.split()])
    return {st: {rk: np.array(rows, dtype=np.int64) for rk, rows in d.items()} for st, d in steps.items()}


def image_grid(images, step):
    """The 4 rank buffers of one image step concatenated in world-rank order (id = rid*2 + cid)."""
    return np.concatenate([images[step][f"P{r // 2}{r % 2}"].ravel() for r in range(4)])


# ---------------------------------------------------------------------------
# natural-layout references (axis 0 = rows gr = y, last axis = columns gc = x)
# ---------------------------------------------------------------------------

def lap_periodic(f):
    """Periodic 5-point (2D) / 7-point (3D) Laplacian, sum of neighbours - 2d * centre."""
    out = -2 * f.ndim * f
    for ax in range(f.ndim):
        out = out + np.roll(f, 1, ax) + np.roll(f, -1, ax)
    return out


def d4_periodic(f, axis, delta=1.0):
    """Periodic 4th-order central difference (f[-2] - 8 f[-1] + 8 f[+1] - f[+2]) / (12 delta)."""
    return (np.roll(f, 2, axis) - 8 * np.roll(f, 1, axis) + 8 * np.roll(f, -1, axis) - np.roll(f, -2, axis)) / (
        12 * delta)


def ramp_dft2(n):
    """Closed-form 2D DFT of the init_host ramp x[gr, gc] = gr*N + gc:
    X[ky, kx] = N^2 R(ky) d(kx) + N d(ky) R(kx), R(0) = N(N-1)/2, R(k) = N / (w^k - 1), w = exp(-2 pi i / N)."""
    k = np.arange(n)
    w = np.exp(-2j * np.pi * k / n)
    R = np.empty(n, dtype=np.complex128)
    R[0] = n * (n - 1) / 2
    R[1:] = n / (w[1:] - 1)
    X = np.zeros((n, n), dtype=np.complex128)
    X[:, 0] += n * n * R
    X[0, :] += n * R
    return X


# ---------------------------------------------------------------------------
# seeded synthetic fields
# ---------------------------------------------------------------------------

def uniform_field(shape, seed):
    """re, im i.i.d. uniform[-1, 1) from numpy's PCG64 default_rng(seed)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


def mode_field(n, dims, nmodes, kmax, seed, noise=0.0):
    """Sum of `nmodes` random Fourier modes with integer wavevectors |k| <= kmax,
    amplitudes N(0,1)/(1+|k|^2), uniform random phases, plus `noise` * U[-1,1)
    in re and im.  Evaluated exactly by placing the modes in a spectrum and
    taking numpy's inverse FFT (scaled so each mode has its amplitude)."""
    rng = np.random.default_rng(seed)
    spec = np.zeros((n,) * dims, dtype=np.complex128)
    got = 0
    while got < nmodes:
        k = rng.integers(-kmax, kmax + 1, size=dims)
        k2 = float(np.sum(k.astype(np.float64) ** 2))
        if k2 > kmax * kmax:
            continue
        amp = rng.normal() / (1.0 + k2)
        ph = rng.uniform(0, 2 * np.pi)
        spec[tuple(int(x) % n for x in k)] += amp * np.exp(1j * ph)
        got += 1
    f = np.fft.ifftn(spec) * (n ** dims)
    if noise:
        f = f + noise * (rng.uniform(-1, 1, f.shape) + 1j * rng.uniform(-1, 1, f.shape))
    return f


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def rel_max(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
