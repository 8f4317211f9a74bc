"""Pins the FP64 CPU oracle before anything is checked against it.

Every check here is independent of the oracle's own code: the reference's
layout images (typed by hand), numpy's FFT, closed forms, the periodic
stencil definitions and their plane-wave eigenvalues, and bijection
invariants of the index maps.  Planted-mistake tests show the checks fail
when a plausible error is introduced into the oracle pipeline.
"""
import itertools

import numpy as np
import pytest

from helpers import (d4_periodic, image_grid, lap_periodic, load_images, ramp_dft2, rel_max, uniform_field)

SHAPES_2D = [(16, 2, 2), (32, 2, 4), (64, 4, 2), (64, 2, 1), (32, 2, 8), (64, 4, 4), (128, 2, 16)]
SHAPES_3D = [(16, 2, 2), (16, 2, 1), (32, 2, 4), (32, 4, 2)]


# ---------------------------------------------------------------------------
# reference fixtures: images/step*.png (N=16, P=2, B=2, do_compute=false)
# ---------------------------------------------------------------------------

def test_images_step0_block_cyclic_is_init_host_before_packing(oracle_mod):
    imgs = load_images()
    s = oracle_mod.Shape(2, 16, 2, 2)
    ramp = s.init_host(np.float64)
    for r in range(4):
        bc = imgs["step0"][f"P{r // 2}{r % 2}"].astype(np.float64).ravel()
        packed = np.empty_like(bc)
        oracle_mod.lib().zo_vector_pack(s.ref, oracle_mod._dp(bc), oracle_mod._dp(packed), 8, 1)
        assert np.array_equal(packed, ramp[r * 64:(r + 1) * 64])


@pytest.mark.parametrize("step,stage", [("step2", 1), ("step3", 2), ("step4", 3), ("step6", 5), ("step7", 7),
                                        ("step8", 8), ("step9", 9)])
def test_images_forward_permutation_stages(oracle_mod, step, stage):
    imgs = load_images()
    s = oracle_mod.Shape(2, 16, 2, 2)
    g = s.init_host()
    assert np.array_equal(image_grid(imgs, "step1"), g.real.astype(np.int64))
    _, stages = s.forward_fft(g, do_compute=False, capture_stages=True)
    assert np.array_equal(image_grid(imgs, step), stages[stage].real.astype(np.int64))


def test_images_inverse_permutation_walks_back(oracle_mod):
    imgs = load_images()
    s = oracle_mod.Shape(2, 16, 2, 2)
    final = image_grid(imgs, "step9").astype(np.complex128)
    back, stages = s.inverse_fft(final, do_compute=False, capture_stages=True)
    # inverse stages: 1 unpack_inverse -> step8, 2 alltoall -> step7, 5 inverse repack -> step4,
    # 7 unpack_inverse -> step3, 8 alltoall -> step2, final -> step1
    assert np.array_equal(stages[1].real.astype(np.int64), image_grid(imgs, "step8"))
    assert np.array_equal(stages[2].real.astype(np.int64), image_grid(imgs, "step7"))
    assert np.array_equal(stages[5].real.astype(np.int64), image_grid(imgs, "step4"))
    assert np.array_equal(stages[7].real.astype(np.int64), image_grid(imgs, "step3"))
    assert np.array_equal(stages[8].real.astype(np.int64), image_grid(imgs, "step2"))
    assert np.array_equal(back.real.astype(np.int64), image_grid(imgs, "step1"))


# ---------------------------------------------------------------------------
# index maps: closed forms and bijections
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n,p,b", SHAPES_2D)
def test_real_maps_bijective_and_closed_form(oracle_mod, n, p, b):
    s = oracle_mod.Shape(2, n, p, b)
    seen = set()
    for rid, cid in itertools.product(range(p), range(p)):
        for lr in range(s.l):
            for lc in range(s.l):
                gr, gc = s.real_to_global(rid, cid, lr, lc)
                # closed form: gr = (lr/B)BP + rid B + lr%B ; gc = jj*M + cid*B + j, lc = j*VEC + jj
                j, jj = divmod(lc, s.vec)
                assert gr == (lr // b) * b * p + rid * b + lr % b
                assert gc == jj * s.m + cid * b + j
                assert s.global_to_real(gr, gc) == (rid, cid, lr, lc)
                seen.add((gr, gc))
    assert len(seen) == n * n


@pytest.mark.parametrize("n,p,b", SHAPES_2D)
def test_reciprocal_maps_bijective(oracle_mod, n, p, b):
    s = oracle_mod.Shape(2, n, p, b)
    seen = set()
    for rid, cid in itertools.product(range(p), range(p)):
        for lr in range(s.l):
            for lc in range(s.l):
                ky, kx = s.reciprocal_to_global(rid, cid, lr, lc)
                assert s.global_to_reciprocal(ky, kx) == (rid, cid, lr, lc)
                seen.add((ky, kx))
    assert len(seen) == n * n
    for i in range(s.l):
        assert s.lp_inv(s.lp(i)) == i
    for q in range(n):
        assert s.ct_inv(s.ct(q)) == q


@pytest.mark.parametrize("n,p,b", SHAPES_3D[:2])
def test_3d_maps_bijective(oracle_mod, n, p, b):
    s = oracle_mod.Shape(3, n, p, b)
    seen_r, seen_k = set(), set()
    for rid, cid, did in itertools.product(range(p), repeat=3):
        for ld, lr, lc in itertools.product(range(s.l), repeat=3):
            g = s.real_to_global_3d(rid, cid, did, ld, lr, lc)
            assert s.global_to_real_3d(*g) == (rid, cid, did, ld, lr, lc)
            k = s.reciprocal_to_global_3d(rid, cid, did, ld, lr, lc)
            assert s.global_to_reciprocal_3d(*k) == (rid, cid, did, ld, lr, lc)
            seen_r.add(g)
            seen_k.add(k)
    assert len(seen_r) == n ** 3 and len(seen_k) == n ** 3


def test_init_host_is_global_linear_index(oracle_mod):
    for dims, (n, p, b) in [(2, (32, 2, 4)), (2, (64, 4, 2)), (3, (16, 2, 2))]:
        s = oracle_mod.Shape(dims, n, p, b)
        nat = s.real_to_natural(s.init_host())
        assert np.array_equal(nat.real.ravel(), np.arange(n ** dims, dtype=np.float64))
        assert not np.any(nat.imag)


# ---------------------------------------------------------------------------
# twiddles: (k, l) index mode and the exp closed form
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n,p,b", [(16, 2, 2), (64, 4, 2)])
def test_twiddle_table(oracle_mod, n, p, b):
    s = oracle_mod.Shape(2, n, p, b)
    for rank in range(p):
        idx = s.twiddle_table(rank, -1, index_mode=True).reshape(s.l, s.l)
        fwd = s.twiddle_table(rank, -1).reshape(s.l, s.l)
        inv = s.twiddle_table(rank, +1).reshape(s.l, s.l)
        for col in range(s.l):
            fr, fc = col % s.vp, col // s.vp
            assert idx[3 % s.l, col] == complex(fr + rank * s.vp, fc)
        k = idx.real
        l_ = idx.imag
        assert np.allclose(fwd, np.exp(-2j * np.pi * k * l_ / n), rtol=0, atol=1e-15)
        assert np.allclose(inv, np.conj(fwd), rtol=0, atol=0)


# ---------------------------------------------------------------------------
# DFT arithmetic
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 128, 3, 12])
def test_dft_matches_numpy(oracle_mod, n):
    x = uniform_field((5, n), 3)
    for mode in (0, 1):
        oracle_mod.set_dft_mode(mode)
        try:
            X = oracle_mod.dft_many(x.ravel(), n, 5, n, 1, -1).reshape(5, n)
            Xi = oracle_mod.dft_many(x.ravel(), n, 5, n, 1, +1).reshape(5, n)
        finally:
            oracle_mod.set_dft_mode(0)
        assert np.allclose(X, np.fft.fft(x, axis=1), rtol=0, atol=1e-13 * n)
        assert np.allclose(Xi, np.fft.ifft(x, axis=1) * n, rtol=0, atol=1e-13 * n)


# ---------------------------------------------------------------------------
# composite transforms
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n,p,b", SHAPES_2D)
def test_forward_2d_is_fft2_through_reciprocal_map(oracle_mod, n, p, b):
    s = oracle_mod.Shape(2, n, p, b)
    nat = uniform_field((n, n), 10 + n + p + b)
    Y = s.reciprocal_to_natural(s.forward_fft(s.natural_to_real(nat)))
    assert rel_max(Y, np.fft.fft2(nat)) < 1e-14


@pytest.mark.parametrize("n,p,b", SHAPES_3D)
def test_forward_3d_is_fftn_through_reciprocal_map(oracle_mod, n, p, b):
    s = oracle_mod.Shape(3, n, p, b)
    nat = uniform_field((n, n, n), 20 + n + p + b)
    Y = s.reciprocal_to_natural(s.forward_fft(s.natural_to_real(nat)))
    assert rel_max(Y, np.fft.fftn(nat)) < 1e-14


@pytest.mark.parametrize("dims,n,p,b", [(2, 32, 2, 4), (2, 64, 4, 2), (3, 16, 2, 2), (3, 32, 2, 4)])
def test_round_trip_heffte_semantics(oracle_mod, dims, n, p, b):
    # heffte/heffte_example_fftw.cpp:105-124: backward(forward(x)) / N^d == x (max-abs check)
    s = oracle_mod.Shape(dims, n, p, b)
    g = s.natural_to_real(uniform_field((n,) * dims, 5))
    back = s.inverse_fft(s.forward_fft(g), scale=1.0 / n ** dims)
    assert np.max(np.abs(back - g)) < 1e-14


def test_ramp_closed_form(oracle_mod):
    for n, p, b in [(32, 2, 4), (64, 4, 2)]:
        s = oracle_mod.Shape(2, n, p, b)
        Y = s.reciprocal_to_natural(s.forward_fft(s.init_host()))
        X = ramp_dft2(n)
        assert rel_max(Y, X) < 1e-14


def test_fast_dft_mode_agrees(oracle_mod):
    s = oracle_mod.Shape(2, 64, 4, 2)
    g = s.natural_to_real(uniform_field((64, 64), 7))
    a = s.forward_fft(g)
    oracle_mod.set_dft_mode(1)
    try:
        b_ = s.forward_fft(g)
    finally:
        oracle_mod.set_dft_mode(0)
    assert rel_max(b_, a) < 1e-14


# ---------------------------------------------------------------------------
# stencils
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n,p,b", SHAPES_2D)
def test_laplace_2d_is_periodic_5_point(oracle_mod, n, p, b):
    s = oracle_mod.Shape(2, n, p, b)
    nat = uniform_field((n, n), 30)
    out = s.real_to_natural(s.laplace(s.natural_to_real(nat)))
    assert np.max(np.abs(out - lap_periodic(nat))) < 1e-14


@pytest.mark.parametrize("n,p,b", SHAPES_3D)
def test_laplace_3d_is_periodic_7_point(oracle_mod, n, p, b):
    s = oracle_mod.Shape(3, n, p, b)
    nat = uniform_field((n, n, n), 31)
    out = s.real_to_natural(s.laplace(s.natural_to_real(nat)))
    assert np.max(np.abs(out - lap_periodic(nat))) < 1e-13


@pytest.mark.parametrize("n,p,b", [(16, 2, 2), (32, 2, 4), (64, 4, 2), (32, 2, 8), (64, 2, 2)])
def test_dx_dy_are_periodic_4th_order(oracle_mod, n, p, b):
    s = oracle_mod.Shape(2, n, p, b)
    rng = np.random.default_rng(40)
    nat = rng.uniform(-1, 1, (n, n))
    g = s.natural_to_real(nat)
    for delta in (1.0, 1.0 / n):
        dx = s.real_to_natural(s.dx(g, delta))
        dy = s.real_to_natural(s.dy(g, delta))
        scale = 1.0 / delta
        assert np.max(np.abs(dx - d4_periodic(nat, 1, delta))) < 1e-14 * scale
        assert np.max(np.abs(dy - d4_periodic(nat, 0, delta))) < 1e-14 * scale
    # complex elements: the real stencil on re and im
    natc = uniform_field((n, n), 41)
    dxc = s.real_to_natural(s.dx(s.natural_to_real(natc)))
    assert np.max(np.abs(dxc - d4_periodic(natc, 1))) < 1e-13


def test_stencils_plane_wave_eigenvalues(oracle_mod):
    n, p, b = 64, 4, 2
    s = oracle_mod.Shape(2, n, p, b)
    gr, gc = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    for kx, ky in [(1, 0), (3, 5), (17, 31), (32, 32)]:
        tx, ty = 2 * np.pi * kx / n, 2 * np.pi * ky / n
        f = np.exp(1j * (tx * gc + ty * gr))
        g = s.natural_to_real(f)
        lap = s.real_to_natural(s.laplace(g))
        assert np.allclose(lap, (2 * np.cos(tx) + 2 * np.cos(ty) - 4) * f, atol=1e-12)
        dx = s.real_to_natural(s.dx(g, 0.5))
        assert np.allclose(dx, 1j * (8 * np.sin(tx) - np.sin(2 * tx)) / (6 * 0.5) * f, atol=1e-12)
        dy = s.real_to_natural(s.dy(g, 0.5))
        assert np.allclose(dy, 1j * (8 * np.sin(ty) - np.sin(2 * ty)) / (6 * 0.5) * f, atol=1e-12)


@pytest.mark.parametrize("n,p,b", [(16, 2, 2), (32, 2, 4), (64, 4, 2), (32, 2, 8)])
def test_deriv4_natural_equals_reference_dx_dy_bitwise(oracle_mod, n, p, b):
    """The natural-layout restatement (used for 3D) reproduces the C restatement
    of dx.cu / dy.cu bit for bit on 2D grids, real and complex."""
    s = oracle_mod.Shape(2, n, p, b)
    rng = np.random.default_rng(42 + n)
    for x in (s.natural_to_real(rng.uniform(-1, 1, (n, n))), s.natural_to_real(uniform_field((n, n), 43 + n))):
        for delta in (1.0, 1.0 / n, 0.3):
            assert np.array_equal(s.deriv4(x, "x", delta).view(np.int64), s.dx(x, delta).view(np.int64))
            assert np.array_equal(s.deriv4(x, "y", delta).view(np.int64), s.dy(x, delta).view(np.int64))


@pytest.mark.parametrize("n,p,b", [(16, 2, 2), (32, 2, 4), (32, 4, 2)])
def test_deriv4_3d_plane_wave_eigenvalues(oracle_mod, n, p, b):
    """dx, dy, dz of exp(i (tx gc + ty gr + tz gd)) = i (8 sin t - sin 2t) / (6 delta) times the wave."""
    s = oracle_mod.Shape(3, n, p, b)
    gd, gr, gc = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    for kx, ky, kz in [(1, 0, 0), (0, 3, 0), (0, 0, 5), (3, 5, 7), (n // 2, 1, n // 2 - 1)]:
        t = {a: 2 * np.pi * k / n for a, k in (("x", kx), ("y", ky), ("z", kz))}
        f = np.exp(1j * (t["x"] * gc + t["y"] * gr + t["z"] * gd))
        g = s.natural_to_real(f)
        for axis in "xyz":
            d = s.real_to_natural(s.deriv4(g, axis, 0.5))
            want = 1j * (8 * np.sin(t[axis]) - np.sin(2 * t[axis])) / (6 * 0.5) * f
            assert np.max(np.abs(d - want)) < 1e-12, (axis, kx, ky, kz)
    # a field constant along z has a zero z derivative (to rounding: the
    # reference's summation order does not cancel exactly) and 2D x / y
    # derivatives per plane
    s2 = oracle_mod.Shape(2, n, p, b)
    plane = uniform_field((n, n), 44)
    vol = np.broadcast_to(plane, (n, n, n)).copy()
    g = s.natural_to_real(vol)
    assert np.max(np.abs(s.deriv4(g, "z"))) < 1e-15
    dx2 = s2.real_to_natural(s2.dx(s2.natural_to_real(plane)))
    assert np.array_equal(s.real_to_natural(s.deriv4(g, "x"))[n // 3], dx2)
    with pytest.raises(ValueError):
        s2.deriv4(s2.natural_to_real(plane), "z")


def test_deriv4_planted_mistakes(oracle_mod):
    """A wrong sign, a swapped axis or a dropped outer term fails the checks above."""
    n, p, b = 16, 2, 2
    s = oracle_mod.Shape(3, n, p, b)
    gd, gr, gc = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    t = 2 * np.pi * 3 / n
    f = np.exp(1j * t * gd)
    want = 1j * (8 * np.sin(t) - np.sin(2 * t)) / 6 * f
    good = oracle_mod.deriv4_natural(f, 0)
    assert np.max(np.abs(good - want)) < 1e-12
    assert np.max(np.abs(-good - want)) > 1e-3                                   # sign
    assert np.max(np.abs(oracle_mod.deriv4_natural(f, 2) - want)) > 1e-3        # axis
    two_point = (np.roll(f, -1, 0) - np.roll(f, 1, 0)) * 8 / 12                   # dropped +-2 terms
    assert np.max(np.abs(two_point - want)) > 1e-3


def test_dx_dy_reject_b1(oracle_mod):
    s = oracle_mod.Shape(2, 64, 2, 1)
    with pytest.raises(ValueError):
        s.dx(np.zeros(64 * 64))


# ---------------------------------------------------------------------------
# planted mistakes: the checks above must fail on plausible oracle errors
# ---------------------------------------------------------------------------

def _pipeline(s, g, *, twiddle=True, swap_axes=False, sign=-1):
    """2D forward built from oracle stages with an optional planted mistake."""
    a = s.fft1a(g, sign)
    a = s.alltoall(1 if swap_axes else 0, s.pack_forward(a))
    a = s.fft1b(s.unpack_forward(1 if swap_axes else 0, a, do_compute=twiddle), sign)
    a = s.repack_transpose(a)
    a = s.fft1a(a, sign)
    a = s.alltoall(0 if swap_axes else 1, s.pack_forward(a))
    return s.fft1b(s.unpack_forward(0 if swap_axes else 1, a, do_compute=twiddle), sign)


def test_planted_mistakes_are_caught(oracle_mod):
    n, p, b = 32, 2, 4
    s = oracle_mod.Shape(2, n, p, b)
    nat = uniform_field((n, n), 50)
    g = s.natural_to_real(nat)
    ref = np.fft.fft2(nat)
    good = s.reciprocal_to_natural(_pipeline(s, g))
    assert rel_max(good, ref) < 1e-14
    for kw in (dict(twiddle=False), dict(swap_axes=True), dict(sign=+1)):
        bad = s.reciprocal_to_natural(_pipeline(s, g, **kw))
        assert rel_max(bad, ref) > 1e-3, kw
    # a wrong stencil sign is caught by the periodic definition
    lap = s.real_to_natural(s.laplace(g))
    assert np.max(np.abs(-lap - lap_periodic(nat))) > 1e-3
