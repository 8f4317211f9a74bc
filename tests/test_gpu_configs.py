"""GPU parity at the exact shapes of the benchmark configurations.

* 2D N=8192, P=4, B=32 (16 virtual ranks) on the seeded smooth field the
  benchmark times (256 random modes with |k| <= N/8, amplitudes
  N(0,1)/(1+|k|^2), uniform phases, plus 1e-3 uniform noise):
  - the 5-point Laplacian and the 4th-order dx/dy in FP64, bit for bit against
    the oracle's restatement of zmodel/stencils/laplace.cu:73-169, dx.cu:10-82
    and dy.cu:10-103; in FP32 within 1e-6 relative L2 of the oracle on the
    same FP32-rounded input;
  - the FP64 forward transform against numpy's pocketfft FFT of the natural
    layout (the CPU library standing in for heFFTe + FFTW,
    heffte/heffte_example_fftw.cpp:78-87), <= 1e-13 relative max error;
  - the FP32 and the split-precision MFMA transforms and round trip against
    the FP64 GPU path on the same rounded input: <= 1e-6 relative L2, relative
    max within the bar of each path.
* 3D 512^3, P=2, B=16 (8 virtual ranks): the FP64 forward / inverse against
  the FP64 library FFT lives in test_gpu_library_fft.py; the 7-point Laplacian
  and dz4 at that shape are checked bit for bit here.
"""
import sys

import numpy as np
import pytest

from conftest import has_gpu
from helpers import MFMA_REL_MAX, FP32_REL_MAX

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]

C2 = (8192, 4, 32)
C2_SEED = 8192
C3 = (512, 2, 16)


def note(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _rel(a, ref):
    import torch
    d = a.to(torch.complex128) - ref.to(torch.complex128)
    return (float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(ref)),
            float(d.abs().max() / ref.abs().max()))


@pytest.fixture(scope="module")
def c2():
    """The C2 field in the packed layout (complex128 on the GPU) and its FP64 plan."""
    from stencilfft_amd import Plan
    from stencilfft_amd.fields import mode_field_natural
    n, p, b = C2
    note(f"C2 field N={n} P={p} B={b} seed={C2_SEED}")
    p64 = Plan(2, n, p, b, precision="fp64")
    nat = mode_field_natural(n, 2, 256, n // 8, C2_SEED, noise=1e-3)
    x = p64.natural_to_packed(nat)
    del nat
    return p64, x


def test_c2_stencils_fp64_bit_exact(oracle_mod, c2):
    p64, x = c2
    n, p, b = C2
    note("C2 FP64 Laplacian / dx4 / dy4 vs oracle, bit for bit")
    s = oracle_mod.Shape(2, n, p, b)
    xh = x.cpu().numpy()
    lap = p64.laplace(x).cpu().numpy()
    assert np.array_equal(lap.view(np.int64), s.laplace(xh).view(np.int64))
    del lap
    for name, want_fn in (("dx4", s.dx), ("dy4", s.dy)):
        got = getattr(p64, name)(x, 1.0 / n).cpu().numpy()
        assert np.array_equal(got.view(np.int64), want_fn(xh, 1.0 / n).view(np.int64)), name
        del got


def test_c2_stencils_fp32(oracle_mod, c2):
    import torch

    from stencilfft_amd import Plan
    p64, x = c2
    n, p, b = C2
    s = oracle_mod.Shape(2, n, p, b)
    p32 = Plan(2, n, p, b, precision="fp32")
    x32 = x.to(torch.complex64)
    xh = x32.to(torch.complex128).cpu().numpy()
    cases = (("laplace", lambda: p32.laplace(x32, scale=float(n * n)), lambda: s.laplace(xh) * float(n * n)),
             ("dx4", lambda: p32.dx4(x32, 1.0 / n), lambda: s.dx(xh, 1.0 / n)),
             ("dy4", lambda: p32.dy4(x32, 1.0 / n), lambda: s.dy(xh, 1.0 / n)))
    for name, gpu, cpu in cases:
        got = gpu()
        want = torch.from_numpy(cpu()).cuda()
        l2, mx = _rel(got, want)
        note(f"C2 FP32 {name} vs oracle: rel L2 {l2:.3e}, rel max {mx:.3e}")
        assert l2 <= 1e-6 and mx <= FP32_REL_MAX, name
        del got, want


def test_c2_fp64_forward_vs_cpu_library(c2):
    """The FP64 GPU path, which every FP32 result is measured against, checked
    against a CPU library FFT of the same field (pocketfft, numpy)."""
    import torch
    p64, x = c2
    n = C2[0]
    note("C2 FP64 forward vs numpy.fft.fft2 on the host")
    y = p64.forward(x)
    tab = p64.reciprocal_index_table()
    nat = torch.empty_like(y)
    nat[tab] = y
    got = nat.cpu().numpy().reshape(n, n)
    del nat, tab, y
    ref = np.fft.fft2(p64.packed_to_natural(x).cpu().numpy().reshape(n, n))
    mx = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    note(f"  rel max {mx:.3e}")
    assert mx <= 1e-13


@pytest.mark.parametrize("precision", ["fp32", "fp32_mfma"])
def test_c2_transforms_vs_fp64_gpu(c2, precision):
    """Forward, and the benchmark's round trip (Laplacian with h = 1/N, forward,
    inverse scaled 1/N^2), against the FP64 GPU path on the same rounded input."""
    import torch

    from stencilfft_amd import Plan
    p64, x = c2
    n, p, b = C2
    note(f"C2 {precision} transforms vs FP64 GPU path")
    pl = Plan(2, n, p, b, precision=precision)
    x32 = x.to(torch.complex64)
    x64 = x32.to(torch.complex128)
    bar = MFMA_REL_MAX if precision == "fp32_mfma" else FP32_REL_MAX
    l2, mx = _rel(pl.forward(x32), p64.forward(x64))
    note(f"  forward rel L2 {l2:.3e}, rel max {mx:.3e}")
    assert l2 <= 1e-6 and mx <= bar
    lap32 = pl.laplace(x32, scale=float(n * n))
    lap64 = p64.laplace(x64, scale=float(n * n))
    z32 = pl.inverse(pl.forward(lap32), 1.0 / float(n) ** 2)
    z64 = p64.inverse(p64.forward(lap64), 1.0 / float(n) ** 2)
    l2, mx = _rel(z32, z64)
    note(f"  round trip rel L2 {l2:.3e}, rel max {mx:.3e}")
    assert l2 <= 1e-6 and mx <= bar


def test_c3_laplace_and_dz4_fp64_bit_exact(oracle_mod):
    """7-point Laplacian and the depth derivative at 512^3 against the oracle."""
    import torch

    from stencilfft_amd import Plan
    n, p, b = C3
    note(f"C3 FP64 Laplacian / dz4 N={n} P={p} B={b}")
    s = oracle_mod.Shape(3, n, p, b)
    pl = Plan(3, n, p, b, precision="fp64")
    g = torch.Generator(device="cuda")
    g.manual_seed(512)
    x = torch.complex(torch.rand(s.total, generator=g, device="cuda", dtype=torch.float64) * 2 - 1,
                      torch.rand(s.total, generator=g, device="cuda", dtype=torch.float64) * 2 - 1)
    xh = x.cpu().numpy()
    got = pl.laplace(x).cpu().numpy()
    assert np.array_equal(got.view(np.int64), s.laplace(xh).view(np.int64))
    del got
    got = pl.dz4(x, 1.0 / n).cpu().numpy()
    assert np.array_equal(got.view(np.int64), s.deriv4(xh, "z", 1.0 / n).view(np.int64))
