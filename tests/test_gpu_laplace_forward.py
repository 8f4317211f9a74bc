"""zm_laplace_forward: the reference's Laplacian followed by forward_fft
(zmodel/main.cu:50-63) as one call.

On FP32 plans with the specialised pass the Laplacian is formed inside the
forward transform's x pass: in 2D for N >= 4096 (rows above and below from
registers, x neighbours from the thread itself and its neighbours' ends), in
3D for lines of N <= 1024 points whose groups of rows lie in one B-block (a
march along the planes: rows above and below and the planes before and after
from registers, x neighbours from the raw tile in LDS).  Elsewhere the call
runs the stencil and then the transform.  Either way the result must be bit for bit
zm_forward(zm_laplace(x, scale)), whose own parity with the oracle
(laplace.cu:73-169, laplace3d.cu:90-225, fft.cu:199-339) the other GPU tests
hold.  Shapes cover both paths, the C2 shape on its smooth mode field, a
batch of two, and the three precisions.
"""
import sys

import numpy as np
import pytest

from conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]

# (dims, n, p, b, fused x pass expected in FP32)
SHAPES = [(2, 256, 2, 8, False), (2, 1024, 2, 16, False), (2, 4096, 4, 16, True), (2, 4096, 2, 16, True),
          (2, 8192, 4, 32, True), (2, 8192, 2, 32, True), (3, 256, 2, 8, False), (3, 512, 2, 16, True),
          (3, 512, 4, 8, True), (3, 512, 2, 8, True), (3, 256, 4, 4, False), (2, 4096, 1, 64, True),
          (3, 256, 1, 16, True)]


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


def _same_bits(a, b):
    import torch
    return bool(torch.equal(torch.view_as_real(a).contiguous().view(torch.int8),
                            torch.view_as_real(b).contiguous().view(torch.int8)))


def _stages(pl, fn):
    import torch
    pl.timing_reset()
    pl.timing(True)
    out = fn()
    pl.timing(False)
    torch.cuda.synchronize()
    st = pl.timing_read()
    pl.timing_reset()
    return out, st


@pytest.mark.parametrize("dims,n,p,b,fused", SHAPES)
@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32_mfma"])
def test_laplace_forward_is_laplace_then_forward(dims, n, p, b, fused, precision):
    import torch
    from stencilfft_amd import Plan
    note(f"laplace_forward {precision} {dims}D N={n} P={p} B={b}")
    pl = Plan(dims, n, p, b, precision=precision)
    x = _field(dims, n, 900 + n + b).to(pl.complex_dtype)
    scale = float(n * n)
    want = pl.forward(pl.laplace(x, scale))
    got, st = _stages(pl, lambda: pl.laplace_forward(x, scale))
    assert _same_bits(got, want)
    expect_fused = fused and precision == "fp32"
    assert ("laplace" not in st) == expect_fused, st
    if expect_fused:
        assert st["line_laplace"][0] == 1 and st["line"][0] == dims - 1 and st["repack"][0] == dims - 1, st
    del torch


def test_c2_smooth_field_and_batch():
    """The benchmark's field (256 smooth modes plus 1e-3 noise) at 8192^2,
    P=4, B=32, in a batch of two with a second field."""
    import torch
    from stencilfft_amd import Plan
    from stencilfft_amd.fields import mode_field_natural
    n, p, b = 8192, 4, 32
    note("laplace_forward C2 smooth field, batch 2")
    p64 = Plan(2, n, p, b, precision="fp64")
    nat = mode_field_natural(n, 2, 256, n // 8, 8192, noise=1e-3)
    x = p64.natural_to_packed(nat).to(torch.complex64)
    del nat, p64
    pl = Plan(2, n, p, b, precision="fp32", batch=2)
    xb = torch.cat([x, _field(2, n, 3).to(torch.complex64)])
    scale = float(n * n)
    assert _same_bits(pl.laplace_forward(xb, scale), pl.forward(pl.laplace(xb, scale)))


def test_3d_batch_of_two():
    """The 3D plane march on a batch of two fields (grid.y = batch): bit for
    bit the two calls, field by field."""
    import torch
    from stencilfft_amd import Plan
    n, p, b = 512, 4, 8
    note(f"laplace_forward 3D N={n} P={p} B={b}, batch 2")
    pl = Plan(3, n, p, b, precision="fp32", batch=2)
    x = _field(3, n, 77, batch=2).to(torch.complex64)
    scale = float(n * n)
    got, st = _stages(pl, lambda: pl.laplace_forward(x, scale))
    assert "line_laplace" in st, st
    assert _same_bits(got, pl.forward(pl.laplace(x, scale)))


def test_3d_lines_of_1024_points():
    """1024^3, P=2, B=32: groups of 4 rows (the only 3D shape class with
    lines of 1024 points), FP32, bit for bit the two calls."""
    import torch
    from stencilfft_amd import Plan
    n, p, b = 1024, 2, 32
    note(f"laplace_forward 3D N={n} P={p} B={b}")
    pl = Plan(3, n, p, b, precision="fp32")
    x = _field(3, n, 78).to(torch.complex64)
    scale = float(n * n)
    got, st = _stages(pl, lambda: pl.laplace_forward(x, scale))
    assert "line_laplace" in st, st
    want = pl.forward(pl.laplace(x, scale))
    assert _same_bits(got, want)


def test_fallback_against_oracle(oracle_mod):
    import torch
    from stencilfft_amd import Plan
    dims, n, p, b = 2, 64, 4, 2
    note(f"laplace_forward vs oracle {dims}D N={n}")
    s = oracle_mod.Shape(dims, n, p, b)
    rng = np.random.default_rng(n)
    x = s.natural_to_real(rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n)))
    pl = Plan(dims, n, p, b, precision="fp64")
    got = pl.laplace_forward(torch.from_numpy(x).cuda(), 3.0).cpu().numpy()
    want = s.forward_fft(s.laplace(x) * 3.0)
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
