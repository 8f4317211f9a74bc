"""zm_forward_noalloc through the C ABI: the forward transform with the
library's static device scratch instead of an allocated one.

Checked against zm_forward with the plan's own scratch (the same kernels, so
bit for bit), on two calls in a row where the second takes a smaller field,
plus the size limit (ZM_ERR_SCRATCH_LIMIT before any launch) and a NULL plan.
"""
import ctypes
import sys

import numpy as np
import pytest

from conftest import has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]


def note(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _field(plan, seed):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    re = torch.rand(plan.numel(), generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    im = torch.rand(plan.numel(), generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    return torch.complex(re, im).to(plan.complex_dtype)


def _both(plan, x):
    """(zm_forward with the plan's scratch, zm_forward_noalloc) of x."""
    import torch
    from stencilfft_amd import _lib as L
    lib = L._lib()
    want = torch.empty_like(x)
    got = torch.full_like(x, float("nan"))
    torch.cuda.synchronize()
    L.check(lib.zm_forward(plan.handle, x.data_ptr(), want.data_ptr(), None, None), "zm_forward")
    L.check(lib.zm_forward_noalloc(plan.handle, x.data_ptr(), got.data_ptr(), None), "zm_forward_noalloc")
    torch.cuda.synchronize()
    return want, got


def _same_bits(a, b):
    import torch
    ra, rb = torch.view_as_real(a), torch.view_as_real(b)
    return bool(torch.equal(ra.contiguous().view(torch.int8), rb.contiguous().view(torch.int8)))


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32_mfma"])
def test_noalloc_matches_forward_on_two_calls(precision):
    from stencilfft_amd import Plan
    big, small = (2, 1024, 2, 16), (2, 256, 4, 8)
    if precision == "fp32_mfma":
        small = (2, 512, 2, 16)
    for dims, n, p, b in (big, small):
        note(f"zm_forward_noalloc {precision} {dims}D N={n} P={p} B={b}")
        pl = Plan(dims, n, p, b, precision=precision)
        x = _field(pl, n)
        want, got = _both(pl, x)
        assert _same_bits(got, want)
        # the result is the transform, not a copy of stale scratch
        assert float(want.abs().max()) > 0


def test_noalloc_3d_and_repeat():
    from stencilfft_amd import Plan
    note("zm_forward_noalloc 3D N=64 P=2 B=4 fp32, twice")
    pl = Plan(3, 64, 2, 4, precision="fp32")
    x = _field(pl, 3)
    w1, g1 = _both(pl, x)
    w2, g2 = _both(pl, x * 2)
    assert _same_bits(g1, w1) and _same_bits(g2, w2)


def test_noalloc_rejects_plans_beyond_the_limit():
    import torch
    from stencilfft_amd import Plan
    from stencilfft_amd import _lib as L
    note("zm_forward_noalloc size limit")
    # 4096^2 complex64: 128 MiB of scratch, twice the static buffer
    pl = Plan(2, 4096, 4, 16, precision="fp32", caller_scratch=True)
    assert pl.scratch_bytes > L.ZM_NOALLOC_SCRATCH_BYTES
    x = torch.zeros(pl.numel(), dtype=torch.complex64, device="cuda")
    y = torch.full_like(x, 7.0)
    st = L._lib().zm_forward_noalloc(pl.handle, x.data_ptr(), y.data_ptr(), None)
    torch.cuda.synchronize()
    assert st == L.ZM_ERR_SCRATCH_LIMIT
    assert bool((y == 7.0).all())  # nothing was launched
    assert L._lib().zm_forward_noalloc(None, x.data_ptr(), y.data_ptr(), None) == L.ZM_ERR_INVALID_ARG
    assert L._lib().zm_forward_noalloc(pl.handle, x.data_ptr(), x.data_ptr(), None) != L.ZM_OK
    msg = L._lib().zm_strerror(L.ZM_ERR_SCRATCH_LIMIT).decode()
    assert "scratch" in msg
