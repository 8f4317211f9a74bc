"""GPU parity: the HIP kernels (through the C ABI) against the FP64 oracle and
the reference's layout fixtures, at sizes the oracle finishes in seconds.

Bars: data movement (permute mode, repack, index maps, layout conversion) is
bit-exact; FP64 stencils are bit-exact (same evaluation order, no FMA
contraction); FP64 transforms within 1e-13 relative max error; FP32 transforms
within the 1e-6 relative-L2 budget of the FP32 path (max error 1e-5).
"""
import ctypes
import sys

import numpy as np
import pytest

from helpers import (FP32_REL_MAX, MFMA_REL_MAX, d4_periodic, image_grid, lap_periodic, load_images, rel_l2, rel_max,
                     uniform_field)

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SHAPES_2D = [(16, 2, 2), (32, 2, 4), (64, 4, 2), (64, 2, 1), (32, 2, 8), (64, 4, 4), (128, 2, 16), (256, 4, 8)]
SHAPES_3D = [(16, 2, 2), (16, 2, 1), (32, 2, 4), (32, 4, 2)]
# VEC >= 16 and L >= 32: the repack's compile-time tiles (two rho per tile at
# VEC = 16, one at VEC = 32; y tiles of 32, 64 and 128)
TILED_SHAPES = [(2, 64, 2, 2), (2, 128, 2, 4), (2, 256, 2, 4), (2, 512, 4, 8), (3, 64, 2, 2), (3, 64, 1, 4),
                (3, 128, 2, 4), (3, 128, 2, 2)]
PATHS = [False, True]  # fused line kernel, two-stage


def note(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def plan(dims, n, p, b, precision="fp64", batch=1, two_stage=False):
    from stencilfft_amd import Plan
    return Plan(dims, n, p, b, precision=precision, batch=batch, two_stage=two_stage)


# ---------------------------------------------------------------------------
# data movement: bit-exact
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("two_stage", PATHS)
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_permute_matches_images(precision, two_stage):
    note(f"permute images {precision} two_stage={two_stage}")
    imgs = load_images()
    pl = plan(2, 16, 2, 2, precision, two_stage=two_stage)
    x = image_grid(imgs, "step1").astype(np.complex128)
    y = host(pl.permute_forward(dev(x, pl.complex_dtype)))
    assert np.array_equal(y.real.astype(np.int64), image_grid(imgs, "step9"))
    back = host(pl.permute_inverse(dev(y, pl.complex_dtype)))
    assert np.array_equal(back.real.astype(np.int64), image_grid(imgs, "step1"))


@pytest.mark.parametrize("two_stage", PATHS)
@pytest.mark.parametrize("dims,n,p,b", [(2,) + s for s in SHAPES_2D] + [(3,) + s for s in SHAPES_3D] + TILED_SHAPES)
def test_permute_bit_exact_vs_oracle(oracle_mod, dims, n, p, b, two_stage):
    note(f"permute {dims}D N={n} P={p} B={b} two_stage={two_stage}")
    s = oracle_mod.Shape(dims, n, p, b)
    rng = np.random.default_rng(n + p + b)
    # random bit patterns in FP64: the kernels must move bits untouched
    bits = rng.integers(0, 2 ** 62, size=2 * s.total, dtype=np.int64).view(np.float64)
    bits = np.where(np.isfinite(bits), bits, 1.0)
    x = bits.view(np.complex128)
    want = s.forward_fft(x, do_compute=False)
    pl = plan(dims, n, p, b, "fp64", two_stage=two_stage)
    got = host(pl.permute_forward(dev(x, torch.complex128)))
    assert np.array_equal(got.view(np.int64), want.view(np.int64))
    back = host(pl.permute_inverse(dev(want, torch.complex128)))
    assert np.array_equal(back.view(np.int64), x.view(np.int64))
    # FP32: integer tags
    pl32 = plan(dims, n, p, b, "fp32", two_stage=two_stage)
    tags = (np.arange(s.total, dtype=np.float64) + 1j * (-np.arange(s.total, dtype=np.float64)))
    got32 = host(pl32.permute_forward(dev(tags, torch.complex64)))
    assert np.array_equal(got32.astype(np.complex128), s.forward_fft(tags, do_compute=False))


@pytest.mark.parametrize("dims,n,p,b", [(2, 64, 4, 2), (2, 32, 2, 8), (3, 16, 2, 2), (3, 32, 2, 4)])
def test_index_tables_and_layout_conversion(oracle_mod, dims, n, p, b):
    note(f"index tables {dims}D N={n}")
    s = oracle_mod.Shape(dims, n, p, b)
    pl = plan(dims, n, p, b, "fp64")
    ramp = s.init_host()
    assert np.array_equal(host(pl.real_index_table()), ramp.real.astype(np.int64))
    lin = np.arange(s.total, dtype=np.float64).astype(np.complex128)
    want_recip = s.natural_to_reciprocal(lin.reshape((n,) * dims))
    assert np.array_equal(host(pl.reciprocal_index_table()), want_recip.real.astype(np.int64))
    nat = uniform_field((n,) * dims, 3)
    packed = host(pl.natural_to_packed(dev(nat.ravel(), torch.complex128)))
    assert np.array_equal(packed, s.natural_to_real(nat))
    back = host(pl.packed_to_natural(dev(packed, torch.complex128)))
    assert np.array_equal(back, nat.ravel())


# ---------------------------------------------------------------------------
# transforms
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("two_stage", PATHS)
@pytest.mark.parametrize("dims,n,p,b", [(2,) + s for s in SHAPES_2D] + [(3,) + s for s in SHAPES_3D])
def test_forward_inverse_fp64_vs_oracle(oracle_mod, dims, n, p, b, two_stage):
    note(f"fp64 fwd/inv {dims}D N={n} P={p} B={b} two_stage={two_stage}")
    s = oracle_mod.Shape(dims, n, p, b)
    x = s.natural_to_real(uniform_field((n,) * dims, 100 + n + p + b))
    pl = plan(dims, n, p, b, "fp64", two_stage=two_stage)
    y = host(pl.forward(dev(x, torch.complex128)))
    want = s.forward_fft(x)
    assert rel_max(y, want) < 1e-13
    scale = 1.0 / n ** dims
    z = host(pl.inverse(dev(want, torch.complex128), scale))
    assert rel_max(z, s.inverse_fft(want, scale=scale)) < 1e-13
    assert np.max(np.abs(z - x)) < 1e-13


@pytest.mark.parametrize("two_stage", PATHS)
@pytest.mark.parametrize("dims,n,p,b", [(2, 64, 4, 2), (2, 256, 4, 8), (2, 128, 2, 16), (3, 32, 2, 4)])
def test_forward_inverse_fp32_vs_oracle(oracle_mod, dims, n, p, b, two_stage):
    note(f"fp32 fwd/inv {dims}D N={n} two_stage={two_stage}")
    s = oracle_mod.Shape(dims, n, p, b)
    x = s.natural_to_real(uniform_field((n,) * dims, 200 + n)).astype(np.complex64).astype(np.complex128)
    pl = plan(dims, n, p, b, "fp32", two_stage=two_stage)
    y = host(pl.forward(dev(x, torch.complex64))).astype(np.complex128)
    want = s.forward_fft(x)
    assert rel_l2(y, want) < 1e-6 and rel_max(y, want) < FP32_REL_MAX
    scale = 1.0 / n ** dims
    w32 = want.astype(np.complex64).astype(np.complex128)
    z = host(pl.inverse(dev(w32, torch.complex64), scale)).astype(np.complex128)
    zwant = s.inverse_fft(w32, scale=scale)
    assert rel_l2(z, zwant) < 1e-6 and rel_max(z, zwant) < FP32_REL_MAX


# shapes whose (VEC, M) pairs run the compile-time specialised line pass
FAST_SHAPES = [(2, 256, 2, 8), (2, 512, 2, 16), (2, 1024, 4, 8), (2, 2048, 2, 16), (2, 4096, 4, 32),
               (2, 2048, 4, 4), (3, 256, 2, 8), (3, 128, 2, 4)]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("dims,n,p,b", FAST_SHAPES)
def test_fast_line_pass_vs_oracle(oracle_mod, dims, n, p, b, precision):
    note(f"fast path {precision} {dims}D N={n} P={p} B={b}")
    s = oracle_mod.Shape(dims, n, p, b)
    rng = np.random.default_rng(n + b)
    x = (rng.uniform(-1, 1, s.total) + 1j * rng.uniform(-1, 1, s.total))
    pl = plan(dims, n, p, b, precision)
    cdt = pl.complex_dtype
    if precision == "fp32":
        x = x.astype(np.complex64).astype(np.complex128)
    oracle_mod.set_dft_mode(1)
    try:
        want = s.forward_fft(x)
        scale = 1.0 / n ** dims
        w_in = want.astype(np.complex64).astype(np.complex128) if precision == "fp32" else want
        zwant = s.inverse_fft(w_in, scale=scale)
    finally:
        oracle_mod.set_dft_mode(0)
    y = host(pl.forward(dev(x, cdt))).astype(np.complex128)
    z = host(pl.inverse(dev(w_in, cdt), scale)).astype(np.complex128)
    if precision == "fp64":
        assert rel_max(y, want) < 1e-13 and rel_max(z, zwant) < 1e-13
    else:
        assert rel_l2(y, want) < 1e-6 and rel_max(y, want) < FP32_REL_MAX
        assert rel_l2(z, zwant) < 1e-6 and rel_max(z, zwant) < FP32_REL_MAX


def test_n1024_p2_b16_fp64_round_trip(oracle_mod):
    # N_DIM=1024, P_DIM=2, B_DIM=16 complex128, re/im U[-1,1): forward, inverse, Laplacian
    note("N=1024 P=2 B=16 fp64")
    n, p, b = 1024, 2, 16
    s = oracle_mod.Shape(2, n, p, b)
    nat = uniform_field((n, n), 1)
    x = s.natural_to_real(nat)
    pl = plan(2, n, p, b, "fp64")
    y = host(pl.forward(dev(x, torch.complex128)))
    oracle_mod.set_dft_mode(1)
    try:
        want = s.forward_fft(x)
    finally:
        oracle_mod.set_dft_mode(0)
    assert rel_max(y, want) < 1e-13
    assert rel_max(s.reciprocal_to_natural(y), np.fft.fft2(nat)) < 1e-13
    z = host(pl.inverse(dev(y, torch.complex128), 1.0 / n ** 2))
    assert np.max(np.abs(z - x)) < 1e-13
    lap = host(pl.laplace(dev(x, torch.complex128)))
    assert np.array_equal(lap.view(np.int64), s.laplace(x).view(np.int64))


def test_batch_fields_are_independent(oracle_mod):
    note("batch of 3 fields")
    n, p, b = 64, 4, 2
    s = oracle_mod.Shape(2, n, p, b)
    xs = [s.natural_to_real(uniform_field((n, n), 300 + k)) for k in range(3)]
    pl = plan(2, n, p, b, "fp64", batch=3)
    y = host(pl.forward(dev(np.concatenate(xs), torch.complex128))).reshape(3, -1)
    for k in range(3):
        assert rel_max(y[k], s.forward_fft(xs[k])) < 1e-13


def test_ramp_closed_form_on_gpu(oracle_mod):
    from helpers import ramp_dft2
    note("ramp closed form")
    for n, p, b in [(64, 4, 2), (256, 4, 8)]:
        s = oracle_mod.Shape(2, n, p, b)
        pl = plan(2, n, p, b, "fp64")
        y = host(pl.forward(dev(s.init_host(), torch.complex128)))
        assert rel_max(s.reciprocal_to_natural(y), ramp_dft2(n)) < 1e-13


# ---------------------------------------------------------------------------
# stencils
# ---------------------------------------------------------------------------

STRIP_SHAPES = [(2, 512, 2, 16), (2, 1024, 4, 8), (2, 1024, 2, 8), (2, 2048, 2, 16), (2, 4096, 2, 32),
                (2, 8192, 4, 64),
                (3, 512, 2, 16)]


@pytest.mark.parametrize("dims,n,p,b", [(2,) + s for s in SHAPES_2D] + [(3,) + s for s in SHAPES_3D] + STRIP_SHAPES)
def test_laplace_fp64_bit_exact(oracle_mod, dims, n, p, b):
    note(f"laplace {dims}D N={n} P={p} B={b}")
    s = oracle_mod.Shape(dims, n, p, b)
    x = s.natural_to_real(uniform_field((n,) * dims, 400 + n))
    pl = plan(dims, n, p, b, "fp64")
    got = host(pl.laplace(dev(x, torch.complex128)))
    assert np.array_equal(got.view(np.int64), s.laplace(x).view(np.int64))
    # real elements and a scale
    xr = np.ascontiguousarray(x.real)
    gotr = host(pl.laplace(dev(xr, torch.float64), scale=float(n * n)))
    assert np.allclose(gotr, s.laplace(xr) * n * n, rtol=1e-15, atol=0)


# 3D shapes of the depth-march Laplacian (row tiles of 8 rows marching over
# depth): L = 32, 64 and 128, one and two x-halo elements per thread, P = 1
# and 2, and the 256^3 P=2 B=8 shape of the batched heat-stepping workload.
DEPTH_MARCH_SHAPES = [(64, 2, 4), (64, 2, 2), (64, 1, 2), (128, 2, 8), (128, 2, 2), (256, 2, 8)]


@pytest.mark.parametrize("n,p,b", DEPTH_MARCH_SHAPES)
def test_laplace_3d_depth_march(oracle_mod, n, p, b):
    """FP64 bit for bit against the oracle (laplace3d.cu:90-225), complex and
    real, a batch of two; FP32 within 1e-6 relative L2."""
    note(f"depth-march laplace 3D N={n} P={p} B={b}")
    s = oracle_mod.Shape(3, n, p, b)
    x = s.natural_to_real(uniform_field((n,) * 3, 700 + n + b))
    x2 = s.natural_to_real(uniform_field((n,) * 3, 800 + n + b))
    want, want2 = s.laplace(x), s.laplace(x2)
    pb = plan(3, n, p, b, "fp64", batch=2)
    got = host(pb.laplace(dev(np.concatenate([x, x2]), torch.complex128)))
    assert np.array_equal(got[:x.size].view(np.int64), want.view(np.int64))
    assert np.array_equal(got[x.size:].view(np.int64), want2.view(np.int64))
    pl = plan(3, n, p, b, "fp64")
    xr = np.ascontiguousarray(x.real)
    gotr = host(pl.laplace(dev(xr, torch.float64), scale=float(n * n)))
    assert np.array_equal(gotr, s.laplace(xr) * float(n * n))
    p32 = plan(3, n, p, b, "fp32")
    g32 = host(p32.laplace(dev(x.astype(np.complex64), torch.complex64))).astype(np.complex128)
    assert rel_l2(g32, s.laplace(x.astype(np.complex64).astype(np.complex128))) < 1e-6


def test_laplace_3d_depth_march_rows_of_512_fp64(oracle_mod):
    """512^3, P=1, B=32: rows of L = 512 packed columns, the depth march's
    tiles of 4 rows (FP64 took the strip kernel at this L before the march's
    halo slots left scratch memory); FP64 bit for bit against the oracle."""
    n, p, b = 512, 1, 32
    note(f"depth-march laplace 3D N={n} P={p} B={b} fp64")
    s = oracle_mod.Shape(3, n, p, b)
    x = uniform_field(n ** 3, 901)  # any field in the packed layout
    pl = plan(3, n, p, b, "fp64")
    got = host(pl.laplace(dev(x, torch.complex128), scale=3.0))
    assert np.array_equal(got.view(np.int64), (s.laplace(x) * 3.0).view(np.int64))


@pytest.mark.parametrize("n,p,b", [(16, 2, 2), (32, 2, 4), (64, 4, 2), (32, 2, 8), (64, 2, 2), (256, 4, 8)])
def test_dx_dy_fp64_bit_exact(oracle_mod, n, p, b):
    note(f"dx/dy N={n} P={p} B={b}")
    s = oracle_mod.Shape(2, n, p, b)
    rng = np.random.default_rng(500 + n)
    xr = s.natural_to_real(rng.uniform(-1, 1, (n, n)))
    xc = s.natural_to_real(uniform_field((n, n), 501 + n))
    pl = plan(2, n, p, b, "fp64")
    for delta in (1.0, 1.0 / n, 0.3):
        for x, dt in ((xr, torch.float64), (xc, torch.complex128)):
            gx = host(pl.dx4(dev(x, dt), delta))
            gy = host(pl.dy4(dev(x, dt), delta))
            assert np.array_equal(gx.view(np.int64), s.dx(x, delta).view(np.int64))
            assert np.array_equal(gy.view(np.int64), s.dy(x, delta).view(np.int64))
            ax = host(pl.dx4(dev(x, dt), delta, adjoint=True))
            assert np.array_equal(ax, -gx)


@pytest.mark.parametrize("n,p,b", [(16, 2, 2), (32, 2, 4), (32, 4, 2), (64, 2, 8), (128, 2, 8)])
def test_dx_dy_dz_3d_fp64_bit_exact(oracle_mod, n, p, b):
    """3D derivatives along x, y and z against the oracle's natural-layout
    restatement (itself bit-identical to dx.cu / dy.cu in 2D)."""
    note(f"3D dx/dy/dz N={n} P={p} B={b}")
    s = oracle_mod.Shape(3, n, p, b)
    rng = np.random.default_rng(510 + n)
    xr = s.natural_to_real(rng.uniform(-1, 1, (n, n, n)))
    xc = s.natural_to_real(uniform_field((n, n, n), 511 + n))
    pl = plan(3, n, p, b, "fp64")
    fns = {"x": pl.dx4, "y": pl.dy4, "z": pl.dz4}
    for delta in (1.0, 1.0 / n, 0.3):
        for x, dt in ((xr, torch.float64), (xc, torch.complex128)):
            for axis, fn in fns.items():
                got = host(fn(dev(x, dt), delta))
                assert np.array_equal(got.view(np.int64), s.deriv4(x, axis, delta).view(np.int64)), axis
                assert np.array_equal(host(fn(dev(x, dt), delta, adjoint=True)), -got)


def test_dz4_needs_3d():
    from stencilfft_amd import ZmError
    note("dz4 on a 2D plan")
    with pytest.raises(ZmError) as e:
        plan(2, 64, 2, 2, "fp64").dz4(torch.zeros(64 * 64, dtype=torch.float64, device="cuda"))
    assert e.value.status == 3
    with pytest.raises(ZmError) as e:
        plan(3, 16, 2, 1, "fp64").dz4(torch.zeros(16 ** 3, dtype=torch.float64, device="cuda"))
    assert e.value.status == 3


@pytest.mark.parametrize("dims,n,p,b", [(2, 256, 4, 8), (3, 32, 2, 4), (3, 512, 2, 16)])
def test_stencils_fp32(oracle_mod, dims, n, p, b):
    note(f"fp32 stencils {dims}D N={n}")
    s = oracle_mod.Shape(dims, n, p, b)
    x = s.natural_to_real(uniform_field((n,) * dims, 600)).astype(np.complex64)
    pl = plan(dims, n, p, b, "fp32")
    got = host(pl.laplace(dev(x, torch.complex64))).astype(np.complex128)
    want = s.laplace(x.astype(np.complex128))
    assert rel_l2(got, want) < 1e-6
    if dims == 2:
        got = host(pl.dx4(dev(x, torch.complex64), 1.0 / n)).astype(np.complex128)
        want = s.dx(x.astype(np.complex128), 1.0 / n)
        assert rel_l2(got, want) < 1e-6
    else:
        # against the FP64 GPU derivatives (bit-exact to the oracle above) on the same rounded input
        p64 = plan(dims, n, p, b, "fp64")
        x32 = dev(x, torch.complex64)
        x64 = x32.to(torch.complex128)
        for f32, f64 in ((pl.dx4, p64.dx4), (pl.dy4, p64.dy4), (pl.dz4, p64.dz4)):
            a = f32(x32, 1.0 / n).to(torch.complex128)
            r = f64(x64, 1.0 / n)
            assert (torch.linalg.vector_norm(a - r) / torch.linalg.vector_norm(r)).item() < 1e-6


def test_stencil_natural_layout_definitions(oracle_mod):
    note("stencils vs periodic definitions")
    n, p, b = 64, 4, 4
    s = oracle_mod.Shape(2, n, p, b)
    nat = uniform_field((n, n), 700)
    pl = plan(2, n, p, b, "fp64")
    x = dev(s.natural_to_real(nat), torch.complex128)
    assert np.max(np.abs(s.real_to_natural(host(pl.laplace(x))) - lap_periodic(nat))) < 1e-13
    assert np.max(np.abs(s.real_to_natural(host(pl.dx4(x, 0.5))) - d4_periodic(nat, 1, 0.5))) < 1e-13
    assert np.max(np.abs(s.real_to_natural(host(pl.dy4(x, 0.5))) - d4_periodic(nat, 0, 0.5))) < 1e-13


# ---------------------------------------------------------------------------
# adjoints and spectral multipliers
# ---------------------------------------------------------------------------

def _dot(a, b):
    return np.vdot(b, a)  # <a, b> = sum conj(b) a


@pytest.mark.parametrize("dims,n,p,b", [(2, 64, 4, 2), (3, 16, 2, 2)])
def test_adjoint_identities(oracle_mod, dims, n, p, b):
    note(f"adjoints {dims}D")
    s = oracle_mod.Shape(dims, n, p, b)
    pl = plan(dims, n, p, b, "fp64")
    x = s.natural_to_real(uniform_field((n,) * dims, 800))
    y = s.natural_to_real(uniform_field((n,) * dims, 801))
    Fx = host(pl.forward(dev(x, torch.complex128)))
    FHy = host(pl.forward_adjoint(dev(y, torch.complex128)))
    assert abs(_dot(Fx, y) - _dot(x, FHy)) < 1e-10 * abs(_dot(Fx, y))
    sc = 0.37
    Ix = host(pl.inverse(dev(x, torch.complex128), sc))
    IHy = host(pl.inverse_adjoint(dev(y, torch.complex128), sc))
    assert abs(_dot(Ix, y) - _dot(x, IHy)) < 1e-10 * abs(_dot(Ix, y))
    Lx = host(pl.laplace(dev(x, torch.complex128)))
    Ly = host(pl.laplace(dev(y, torch.complex128)))
    assert abs(_dot(Lx, y) - _dot(x, Ly)) < 1e-10 * abs(_dot(Lx, y))
    for name in ("dx4", "dy4", "dz4") if dims == 3 else ("dx4", "dy4"):
        fn = getattr(pl, name)
        D = host(fn(dev(x, torch.complex128), 0.25))
        DH = host(fn(dev(y, torch.complex128), 0.25, adjoint=True))
        assert abs(_dot(D, y) - _dot(x, DH)) < 1e-10 * abs(_dot(D, y)), name


@pytest.mark.parametrize("dims,n,p,b", [(2, 64, 4, 2), (3, 32, 2, 4)])
def test_spectral_multipliers(oracle_mod, dims, n, p, b):
    note(f"spectral {dims}D")
    s = oracle_mod.Shape(dims, n, p, b)
    pl = plan(dims, n, p, b, "fp64")
    x = uniform_field(s.total, 900).astype(np.complex128)
    k = np.fft.fftfreq(n, 1.0 / n)
    grids = np.meshgrid(*([k] * dims), indexing="ij")
    k2 = sum(g ** 2 for g in grids) * (2 * np.pi) ** 2
    k2_grid = s.natural_to_reciprocal(k2.astype(np.complex128)).real
    pois = np.where(k2_grid == 0, 0.0, -1.0 / np.where(k2_grid == 0, 1.0, k2_grid))
    got = host(pl.spectral_multiply(dev(x, torch.complex128), "poisson"))
    assert np.allclose(got, x * pois, rtol=1e-15, atol=0)
    got = host(pl.spectral_multiply(dev(x, torch.complex128), "heat", 1e-3, 1e-2))
    assert np.allclose(got, x * np.exp(-1e-3 * k2_grid * 1e-2), rtol=1e-14, atol=0)


def test_poisson_solve_recovers_laplacian_inverse(oracle_mod):
    """u = IFFT(FFT(f) / (-|k|^2)) solves the spectral Poisson equation: FFT(u) * (-|k|^2) = FFT(f) - mean."""
    note("poisson solve")
    n, p, b = 64, 4, 2
    s = oracle_mod.Shape(2, n, p, b)
    pl = plan(2, n, p, b, "fp64")
    gr, gc = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    f = np.sin(2 * np.pi * 3 * gc / n) * np.cos(2 * np.pi * 5 * gr / n) + 0j
    x = dev(s.natural_to_real(f), torch.complex128)
    u = pl.inverse(pl.spectral_multiply(pl.forward(x), "poisson"), 1.0 / n ** 2)
    un = s.real_to_natural(host(u))
    want = -f / ((2 * np.pi) ** 2 * (3 ** 2 + 5 ** 2))
    assert np.max(np.abs(un - want)) < 1e-15


# ---------------------------------------------------------------------------
# error paths
# shapes with radix-16 passes in both the VEC- and the M-point DFTs, where the
# split-precision MFMA variant replaces them (VEC, M in {16, ..., 128})
MFMA_SHAPES = [(2, 256, 2, 8), (2, 512, 2, 16), (2, 1024, 4, 8), (2, 2048, 2, 16), (2, 4096, 4, 32),
               (2, 2048, 4, 4), (3, 256, 2, 8), (3, 256, 4, 4)]


@pytest.mark.parametrize("dims,n,p,b", MFMA_SHAPES)
def test_mfma_line_pass_vs_oracle(oracle_mod, dims, n, p, b):
    """FP32 data with the radix-16 DFT stages as split-precision (hi/lo FP16,
    FP32-accumulated) MFMA GEMMs: same bar as the FP32 vector-ALU path."""
    note(f"fp32_mfma {dims}D N={n} P={p} B={b}")
    s = oracle_mod.Shape(dims, n, p, b)
    rng = np.random.default_rng(3 * n + b)
    x = (rng.uniform(-1, 1, s.total) + 1j * rng.uniform(-1, 1, s.total)).astype(np.complex64).astype(np.complex128)
    pl = plan(dims, n, p, b, "fp32_mfma")
    oracle_mod.set_dft_mode(1)
    try:
        want = s.forward_fft(x)
        scale = 1.0 / n ** dims
        w_in = want.astype(np.complex64).astype(np.complex128)
        zwant = s.inverse_fft(w_in, scale=scale)
    finally:
        oracle_mod.set_dft_mode(0)
    y = host(pl.forward(dev(x, torch.complex64))).astype(np.complex128)
    z = host(pl.inverse(dev(w_in, torch.complex64), scale)).astype(np.complex128)
    note(f"  forward rel L2 {rel_l2(y, want):.2e} max {rel_max(y, want):.2e}; "
         f"inverse rel L2 {rel_l2(z, zwant):.2e} max {rel_max(z, zwant):.2e}")
    assert rel_l2(y, want) < 1e-6 and rel_max(y, want) < MFMA_REL_MAX
    assert rel_l2(z, zwant) < 1e-6 and rel_max(z, zwant) < MFMA_REL_MAX


def test_mfma_spiky_spectrum():
    """Per-transform power-of-two scaling: a field of a few huge modes plus
    tiny noise (a spectrum spanning ~12 decades) keeps the FP32 error bar."""
    from stencilfft_amd.fields import mode_fields
    note("fp32_mfma spiky spectrum")
    n, p, b = 2048, 2, 16
    p64 = plan(2, n, p, b, "fp64")
    pm = plan(2, n, p, b, "fp32_mfma")
    x64 = mode_fields(p64, 4, 8, seed=5) * 1e6 + mode_fields(p64, 16, n // 4, seed=6) * 1e-6
    x32 = x64.to(torch.complex64)
    x64r = x32.to(torch.complex128)
    y64 = p64.forward(x64r)
    ym = pm.forward(x32).to(torch.complex128)
    rel = (torch.linalg.vector_norm(ym - y64) / torch.linalg.vector_norm(y64)).item()
    z64 = p64.inverse(y64.to(torch.complex64).to(torch.complex128), 1.0 / n ** 2)
    zm = pm.inverse(y64.to(torch.complex64), 1.0 / n ** 2).to(torch.complex128)
    relz = (torch.linalg.vector_norm(zm - z64) / torch.linalg.vector_norm(z64)).item()
    note(f"  forward rel L2 {rel:.2e}, inverse {relz:.2e}")
    assert rel < 1e-6 and relz < 1e-6


# ---------------------------------------------------------------------------

def test_error_paths():
    from stencilfft_amd import ZmError
    note("error paths")
    pl = plan(2, 64, 2, 2, "fp64")
    x = torch.zeros(64 * 64, dtype=torch.complex128, device="cuda")
    with pytest.raises(ZmError) as e:
        pl.forward(x, out=x)
    assert e.value.status == 5
    with pytest.raises(ValueError):
        pl.forward(torch.zeros(64 * 64, dtype=torch.complex64, device="cuda"))
    p1 = plan(2, 64, 2, 1, "fp64")
    with pytest.raises(ZmError) as e:
        p1.dx4(torch.zeros(64 * 64, dtype=torch.float64, device="cuda"))
    assert e.value.status == 3
    with pytest.raises(ZmError) as e:
        plan(2, 64, 2, 2, "fp32_mfma")  # M = 4: no specialised line pass
    assert e.value.status == 3
    with pytest.raises(ZmError) as e:
        plan(2, 1024, 2, 16, "fp32_mfma", two_stage=True)
    assert e.value.status == 3
    # The C ABI with a live plan: NULL buffers, a scratch buffer that aliases
    # an operand, an unknown element kind and a zero or NaN step are rejected
    # before any launch.
    from stencilfft_amd import _lib as L
    lib, h = L._lib(), pl.handle
    y = torch.zeros_like(x)
    xp, yp = ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr())
    assert lib.zm_forward(h, None, yp, None, None) == L.ZM_ERR_INVALID_ARG
    assert lib.zm_inverse(h, xp, None, 1.0, None, None) == L.ZM_ERR_INVALID_ARG
    assert lib.zm_forward(h, xp, yp, xp, None) == L.ZM_ERR_ALIAS
    assert lib.zm_laplace(h, xp, None, 1.0, 0, None) == L.ZM_ERR_INVALID_ARG
    assert lib.zm_laplace(h, xp, yp, 1.0, 7, None) == L.ZM_ERR_INVALID_ARG
    assert lib.zm_dx4(h, xp, yp, 0.0, 0, None) == L.ZM_ERR_INVALID_ARG
    assert lib.zm_dy4(h, xp, yp, float("nan"), 0, None) == L.ZM_ERR_INVALID_ARG
    assert lib.zm_natural_to_packed(h, None, yp, None) == L.ZM_ERR_INVALID_ARG
    torch.cuda.synchronize()


@pytest.mark.parametrize("n,p,b,kmax", [(2048, 4, 8, 16), (4096, 4, 16, 512)])
def test_fp32_stencils_accurate_on_smooth_fields(n, p, b, kmax):
    """Smooth fields make sum-of-neighbours minus centre cancel; the FP32
    stencils must still agree with FP64 on the same rounded input."""
    from stencilfft_amd.fields import mode_fields
    note(f"fp32 stencils on smooth fields N={n} kmax={kmax}")
    p32 = plan(2, n, p, b, "fp32")
    p64 = plan(2, n, p, b, "fp64")
    x64 = mode_fields(p64, 64, kmax, seed=41)
    x32 = x64.to(torch.complex64)
    x64r = x32.to(torch.complex128)

    def rel(a, r):
        return (torch.linalg.vector_norm(a.to(torch.complex128) - r) / torch.linalg.vector_norm(r)).item()

    assert rel(p32.laplace(x32, float(n * n)), p64.laplace(x64r, float(n * n))) < 1e-6
    assert rel(p32.dx4(x32, 1.0 / n), p64.dx4(x64r, 1.0 / n)) < 1e-6
    assert rel(p32.dy4(x32, 1.0 / n), p64.dy4(x64r, 1.0 / n)) < 1e-6


@pytest.mark.parametrize("dims,n,p,b,precision", [(2, 1024, 2, 16, "fp32"), (2, 256, 4, 8, "fp64"),
                                                  (3, 64, 2, 4, "fp32")])
def test_graph_capture_replay(dims, n, p, b, precision):
    """Every launch is stream-ordered on the caller's stream and the library
    records no memset, so a Laplacian -> forward -> inverse round trip can be
    captured in a HIP graph.  Each replay (with eager work on the same buffers
    between replays) must equal the eager result bit for bit, also after the
    input changes."""
    from stencilfft_amd.fields import uniform_fields
    note(f"graph capture {dims}D N={n} {precision}")
    pl = plan(dims, n, p, b, precision)
    x = uniform_fields(pl, 7)
    lap, y, z = torch.empty_like(x), torch.empty_like(x), torch.empty_like(x)
    scratch = torch.empty(pl.scratch_bytes // x.element_size(), dtype=x.dtype, device="cuda")
    scale = 1.0 / float(n) ** dims

    def run():
        pl.laplace(x, float(n * n), out=lap)
        pl.forward(lap, out=y, scratch=scratch)
        pl.inverse(y, scale, out=z, scratch=scratch)

    def eager(inp):
        zz = torch.empty_like(inp)
        pl.inverse(pl.forward(pl.laplace(inp, float(n * n))), scale, out=zz)
        torch.cuda.synchronize()
        return zz

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        run()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        run()
    for i in range(4):
        if i == 2:
            x.copy_(uniform_fields(pl, 8))
        want = eager(x)
        z.zero_()
        lap.mul_(0)  # eager work on the graph's buffers between replays
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(z, want), f"replay {i}"


def test_plan_path_query():
    note("plan path")
    assert plan(2, 8192, 4, 32, "fp32").fast
    assert plan(2, 1024, 2, 16, "fp32_mfma").mfma
    p2 = plan(2, 1024, 2, 16, "fp64", two_stage=True)
    assert p2.two_stage and not p2.fast
    p3 = plan(2, 64, 2, 2, "fp32")  # M = 4: generic line pass
    assert not p3.fast and not p3.two_stage


SPECIALISED_PAIRS = [(v, m) for v in (16, 32, 64, 128) for m in (16, 32, 64, 128) if v * m <= 8192]


@pytest.mark.parametrize("precision", ["fp32", "fp64", "fp32_mfma"])
def test_every_specialised_pair_takes_the_fused_pass(precision):
    # N = VEC * M with P = 2 and B = M / 2 gives exactly that (VEC, M); every
    # pair the specialised pass is built for must be selected, in each
    # precision (a dispatch gap would fall back to the generic pass silently)
    note(f"fused pass for every (VEC, M) pair, {precision}")
    for vec, m in SPECIALISED_PAIRS:
        p = plan(2, vec * m, 2, m // 2, precision)
        assert p.fast, (vec, m)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_concurrent_streams_with_own_scratch(precision):
    """Calls on one plan from two streams at once, each with its own scratch
    buffer, give the same bits as the same calls made one after the other."""
    from stencilfft_amd.fields import uniform_fields
    note(f"two streams {precision}")
    pl = plan(2, 1024, 2, 16, precision)
    xs = [uniform_fields(pl, 70 + i) for i in range(2)]
    want = []
    for x in xs:
        y = pl.forward(x)
        want.append((y, pl.inverse(y, 1.0 / 1024 ** 2)))
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    scratch = [torch.empty(pl.scratch_bytes // xs[0].element_size(), dtype=xs[0].dtype, device="cuda")
               for _ in range(2)]
    outs = [(torch.empty_like(x), torch.empty_like(x)) for x in xs]
    for s in streams:
        s.wait_stream(torch.cuda.current_stream())
    for _ in range(5):  # repeated so that the two streams' kernels overlap
        for i, s in enumerate(streams):
            pl.forward(xs[i], out=outs[i][0], scratch=scratch[i], stream=s)
            pl.inverse(outs[i][0], 1.0 / 1024 ** 2, out=outs[i][1], scratch=scratch[i], stream=s)
    torch.cuda.synchronize()
    for (y, z), (wy, wz) in zip(outs, want):
        assert torch.equal(y, wy) and torch.equal(z, wz)


def test_scratch_checked_before_launch():
    """A caller's scratch must be a contiguous device tensor of at least
    scratch_bytes that overlaps neither field; anything else raises before a
    kernel is launched.  A caller-scratch plan owns none and refuses NULL."""
    from stencilfft_amd import ZmError
    from stencilfft_amd import _lib as L
    note("scratch checks")
    pl = plan(2, 256, 4, 8, "fp32")
    x = torch.zeros(256 * 256, dtype=torch.complex64, device="cuda")
    y = torch.full_like(x, 3.0)
    elems = pl.scratch_bytes // x.element_size()
    bad = [torch.empty(elems - 1, dtype=x.dtype, device="cuda"),           # one element short
           torch.empty(elems, dtype=x.dtype),                              # host memory
           torch.empty(2 * elems, dtype=x.dtype, device="cuda")[::2]]      # not contiguous
    for s in bad:
        with pytest.raises(ValueError):
            pl.forward(x, out=y, scratch=s)
    big = torch.empty(2 * elems, dtype=x.dtype, device="cuda")
    with pytest.raises(ValueError):
        pl.forward(x, out=big[:elems], scratch=big[elems // 2:])  # scratch overlaps the output
    torch.cuda.synchronize()
    assert torch.all(y == 3.0)  # nothing ran
    from stencilfft_amd import Plan
    pc = Plan(2, 256, 4, 8, precision="fp32", caller_scratch=True)
    with pytest.raises(ValueError):
        pc.forward(x, out=y)
    lib = L._lib()
    xp, yp = ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr())
    assert lib.zm_forward(pc.handle, xp, yp, None, None) == L.ZM_ERR_INVALID_ARG
    s = torch.empty(elems, dtype=x.dtype, device="cuda")
    assert torch.equal(pc.forward(x, scratch=s), pl.forward(x))
    with pytest.raises(ZmError) as e:
        L.check(lib.zm_plan_create(ctypes.byref(ctypes.c_void_p()), ctypes.byref(L.ZmShape(2, 256, 4, 8)),
                                   L.ZM_FP32, 1, 1 << 7), "zm_plan_create")
    assert e.value.status == L.ZM_ERR_INVALID_ARG


def test_mfma_tiny_magnitudes():
    """Line groups whose largest value is an FP32 denormal or a tiny normal
    keep a finite power-of-two scale and the FP32 error bar."""
    note("fp32_mfma tiny magnitudes")
    n, p, b = 512, 2, 16
    pm = plan(2, n, p, b, "fp32_mfma")
    p64 = plan(2, n, p, b, "fp64")
    g = torch.Generator(device="cuda")
    g.manual_seed(13)
    re = torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    im = torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    for scale in (2.0 ** -130, 2.0 ** -118):
        x32 = (torch.complex(re, im) * scale).to(torch.complex64)
        x64 = x32.to(torch.complex128)
        y = pm.forward(x32).to(torch.complex128)
        assert torch.isfinite(y).all()
        r = p64.forward(x64)
        rel = (torch.linalg.vector_norm(y - r) / torch.linalg.vector_norm(r)).item()
        note(f"  scale 2^{int(np.log2(scale))}: forward rel L2 {rel:.2e}")
        assert rel < 1e-5  # denormal inputs carry fewer than 24 significant bits
