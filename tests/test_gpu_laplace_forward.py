"""zm_laplace_forward: the reference's Laplacian followed by forward_fft
(zmodel/main.cu:50-63) as one chain, the Laplacian formed inside the last
direction's line pass.

* FP64 against the two-call sequence zm_laplace + zm_forward on the GPU, and
  against the oracle's restatement of laplace.cu:73-169 / laplace3d.cu:90-225
  followed by fft.cu:199-339 / fft3d.cu:234-412: <= 1e-13 relative max error
  (the fused form is not bit-identical: it applies the second difference of
  the transformed directions as its exact symbol -4 sin^2(pi k / N)).
* FP32 and the split-precision MFMA path against the FP64 two-call sequence on
  the same FP32-rounded input: <= 1e-6 relative L2, <= 1e-5 relative max.
* Shapes cover every (VEC, M) family of the specialised pass that the 2D and
  3D benchmark shapes use, a batch of two, and a plan without the
  specialised pass (the two-call fallback).
"""
import sys

import numpy as np
import pytest

from conftest import has_gpu
from helpers import FP32_REL_MAX, MFMA_REL_MAX

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]

SHAPES = [(2, 256, 2, 8), (2, 1024, 2, 16), (2, 2048, 4, 8), (2, 4096, 4, 16), (2, 8192, 4, 32), (2, 8192, 2, 32),
          (3, 256, 2, 8), (3, 512, 2, 16)]


def note(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _field(dims, n, seed, batch=1):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    m = batch * n ** dims
    re = torch.rand(m, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    im = torch.rand(m, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    return torch.complex(re, im)


def _rel(a, ref):
    import torch
    d = a.to(torch.complex128) - ref.to(torch.complex128)
    return (float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(ref)),
            float(d.abs().max() / ref.abs().max()))


@pytest.mark.parametrize("dims,n,p,b", SHAPES)
def test_fused_matches_two_calls(dims, n, p, b):
    from stencilfft_amd import Plan
    note(f"laplace_forward {dims}D N={n} P={p} B={b}")
    scale = float(n * n)
    p64 = Plan(dims, n, p, b, precision="fp64")
    x = _field(dims, n, 900 + n + b)
    want = p64.forward(p64.laplace(x, scale))
    got = p64.laplace_forward(x, scale)
    l2, mx = _rel(got, want)
    note(f"  fp64 vs laplace + forward: rel L2 {l2:.2e}, rel max {mx:.2e}")
    assert mx < 1e-13
    x32 = x.to(__import__("torch").complex64)
    want32 = p64.forward(p64.laplace(x32.to(x.dtype), scale))
    for prec, bar in (("fp32", FP32_REL_MAX), ("fp32_mfma", MFMA_REL_MAX)):
        pl = Plan(dims, n, p, b, precision=prec)
        l2, mx = _rel(pl.laplace_forward(x32, scale), want32)
        note(f"  {prec} vs fp64: rel L2 {l2:.2e}, rel max {mx:.2e}")
        assert l2 <= 1e-6 and mx <= bar
        # the same against the two-call FP32 path, whose error it should match
        l2b, _ = _rel(pl.forward(pl.laplace(x32, scale)), want32)
        assert l2 <= 2.0 * l2b + 1e-7


def test_fused_batch_and_round_trip():
    import torch
    from stencilfft_amd import Plan
    dims, n, p, b = 2, 2048, 4, 8
    note(f"laplace_forward batch 2, round trip {dims}D N={n}")
    scale = float(n * n)
    pl = Plan(dims, n, p, b, precision="fp64", batch=2)
    x = _field(dims, n, 77, batch=2)
    got = pl.laplace_forward(x, scale)
    want = pl.forward(pl.laplace(x, scale))
    assert _rel(got, want)[1] < 1e-13
    z = pl.inverse(got, 1.0 / n ** 2)
    assert _rel(z, pl.laplace(x, scale))[1] < 1e-13
    del torch


@pytest.mark.parametrize("dims,n,p,b", [(2, 64, 4, 2), (3, 32, 2, 4)])
def test_fused_against_oracle_and_fallback(oracle_mod, dims, n, p, b):
    """Plans without the specialised pass take the two-call fallback; both
    paths against the oracle's Laplacian + forward transform."""
    import torch
    from stencilfft_amd import Plan
    note(f"laplace_forward vs oracle {dims}D N={n} P={p} B={b}")
    s = oracle_mod.Shape(dims, n, p, b)
    rng = np.random.default_rng(n)
    nat = rng.uniform(-1, 1, (n,) * dims) + 1j * rng.uniform(-1, 1, (n,) * dims)
    x = s.natural_to_real(nat)
    want = s.forward_fft(s.laplace(x) * 3.0)
    pl = Plan(dims, n, p, b, precision="fp64")
    got = pl.laplace_forward(torch.from_numpy(x).cuda(), 3.0).cpu().numpy()
    assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-13


def test_fused_against_oracle_on_the_fast_path(oracle_mod):
    import torch
    from stencilfft_amd import Plan
    dims, n, p, b = 2, 256, 2, 8
    note(f"laplace_forward vs oracle (specialised pass) {dims}D N={n}")
    s = oracle_mod.Shape(dims, n, p, b)
    rng = np.random.default_rng(5)
    x = s.natural_to_real(rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n)))
    pl = Plan(dims, n, p, b, precision="fp64")
    from stencilfft_amd import _lib as L
    path = __import__("ctypes").c_uint()
    L.check(L._lib().zm_plan_path(pl.handle, __import__("ctypes").byref(path)))
    assert path.value & L.ZM_PATH_FAST_LINE
    got = pl.laplace_forward(torch.from_numpy(x).cuda(), float(n * n)).cpu().numpy()
    want = s.forward_fft(s.laplace(x) * float(n * n))
    assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-13


def test_two_stage_plans_are_unsupported():
    import torch
    from stencilfft_amd import Plan, ZmError
    note("laplace_forward on a two-stage plan")
    pl = Plan(2, 256, 2, 8, precision="fp64", two_stage=True)
    x = torch.zeros(pl.numel(), dtype=torch.complex128, device="cuda")
    with pytest.raises(ZmError) as e:
        pl.laplace_forward(x, 1.0)
    assert e.value.status == 3
