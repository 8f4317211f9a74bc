"""The packed transforms against a library FFT on the natural layout.

torch.fft.fftn / ifftn (hipFFT over rocFFT on this platform) plays the role of
the reference's heFFTe + cuFFT comparison (heffte/heffte_example_cuda.cpp,
heffte_example_cuda3d.cpp): the same field, distributed as pencils there and
held whole here, transformed by a vendor library.  Our forward output is
scattered to natural order with the reciprocal index table
(zmodel/utils.h:254-320); the inverse takes the packed spectrum back.

Bars: FP64 <= 1e-13 relative to the largest output; FP32 <= 1e-6 relative L2
against the FP64 library result, and relative max error within the bar of the
FP32 path (FP32_REL_MAX, MFMA_REL_MAX in helpers.py); the library's own FP32
error is printed for comparison.
"""
import pytest

from conftest import has_gpu
from helpers import FP32_REL_MAX, MFMA_REL_MAX

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]

CASES = [
    (2, 1024, 2, 16, "fp64"), (2, 1024, 2, 16, "fp32"), (2, 8192, 4, 32, "fp32"), (2, 4096, 4, 16, "fp64"),
    (3, 128, 2, 8, "fp64"), (3, 512, 2, 16, "fp32"), (2, 2048, 2, 16, "fp32_mfma"), (2, 8192, 4, 32, "fp32_mfma"),
    # FP64 line passes of N = 8192 in both (VEC, M) orders: (64, 128) and (128, 64)
    (2, 8192, 4, 32, "fp64"), (2, 8192, 4, 16, "fp64"),
    # the 3D configuration, 512^3 P=2 B=16, in FP64
    (3, 512, 2, 16, "fp64"),
]


def _natural_field(torch, dims, n, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    shape = (n,) * dims
    re = torch.rand(shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    im = torch.rand(shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    return torch.complex(re, im)


def _rel(a, ref, torch):
    d = a.to(torch.complex128) - ref
    return (float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(ref)),
            float(d.abs().max() / ref.abs().max()))


@pytest.mark.parametrize("dims,n,p,b,precision", CASES)
def test_forward_inverse_match_library_fft(dims, n, p, b, precision):
    import torch

    from stencilfft_amd import Plan
    print(f"\ncase {dims}D N={n} P={p} B={b} {precision}", flush=True)
    plan = Plan(dims, n, p, b, precision=precision)
    nat64 = _natural_field(torch, dims, n, seed=n + dims)
    ref = torch.fft.fftn(nat64, dim=tuple(range(dims))).reshape(-1)
    nat = nat64.to(plan.complex_dtype).reshape(-1).contiguous()
    y = plan.forward(plan.natural_to_packed(nat))
    tab = plan.reciprocal_index_table()
    got = torch.empty_like(y)
    got[tab] = y
    l2, mx = _rel(got, ref, torch)
    lib32 = None
    if precision != "fp64":
        lib = torch.fft.fftn(nat.reshape((n,) * dims), dim=tuple(range(dims))).reshape(-1)
        lib32 = _rel(lib, ref, torch)
    print(f"forward vs FP64 library: rel L2 {l2:.3e}, rel max {mx:.3e}; FP32 library {lib32}", flush=True)
    bar = MFMA_REL_MAX if precision == "fp32_mfma" else FP32_REL_MAX
    if precision == "fp64":
        assert mx <= 1e-13
    else:
        assert l2 <= 1e-6 and mx <= bar
    # inverse of the packed spectrum of the library transform, scaled 1/N^d
    spec = ref.to(plan.complex_dtype)[tab].contiguous()
    back = plan.packed_to_natural(plan.inverse(spec, 1.0 / float(n) ** dims))
    l2i, mxi = _rel(back, nat64.reshape(-1), torch)
    print(f"inverse vs input: rel L2 {l2i:.3e}, rel max {mxi:.3e}", flush=True)
    if precision == "fp64":
        assert mxi <= 1e-13
    else:
        assert l2i <= 1e-6 and mxi <= bar
