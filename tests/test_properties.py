"""Property tests over random valid (dims, N, P, B) on the CPU (hypothesis).

The C ABI's derived sizes and host index maps against their closed forms
(zmodel/utils.h:28-71, 122-320) and the oracle; the oracle's transforms
against numpy's FFT through the reciprocal map.  Shapes are drawn as
log2 P, log2 B and log2 VP (VEC = P * VP), so every draw is a valid
interleaved block-cyclic distribution, P = 1 and P = 8 included.
"""
import numpy as np
from hypothesis import HealthCheck, given, settings, strategies as st

from helpers import rel_max, uniform_field


@st.composite
def shapes(draw, dims=None, max_log2n=(8, 5)):
    d = dims if dims is not None else draw(st.sampled_from([2, 3]))
    lp = draw(st.integers(0, 3))
    lb = draw(st.integers(0, 4))
    lvp = draw(st.integers(0, 3))
    ln = 2 * lp + lb + lvp  # N = P * B * VEC, VEC = P * VP
    cap = max_log2n[0] if d == 2 else max_log2n[1]
    if ln > cap:  # shrink B, then VP, to fit the size cap
        cut = min(lb, ln - cap)
        lb -= cut
        ln -= cut
        cut = min(lvp, ln - cap)
        lvp -= cut
        ln -= cut
    if ln > cap:
        lp = max(0, lp - (ln - cap + 1) // 2)
        ln = 2 * lp + lb + lvp
    return d, 1 << ln, 1 << lp, 1 << lb


SETTINGS = dict(max_examples=60, deadline=None, derandomize=True, suppress_health_check=[HealthCheck.too_slow])


@settings(**SETTINGS)
@given(shapes())
def test_shape_check_closed_forms(shape):
    from stencilfft_amd import Shape
    dims, n, p, b = shape
    d = Shape(dims, n, p, b).check()
    assert d.l * p == n and d.vec * b == d.l and d.vp * p == d.vec and d.m == b * p
    assert d.nranks == p ** dims and d.local == d.l ** dims and d.field == n ** dims
    assert d.vec * d.m == n  # the two DFT lengths of one line


@settings(**SETTINGS)
@given(shapes(max_log2n=(7, 4)), st.integers(0, 2 ** 31 - 1))
def test_host_maps_round_trip_and_match_oracle(oracle_mod, shape, seed):
    from stencilfft_amd import Shape
    dims, n, p, b = shape
    s = Shape(dims, n, p, b)
    o = oracle_mod.Shape(dims, n, p, b)
    rng = np.random.default_rng(seed)
    l = n // p
    for _ in range(20):
        coords = tuple(int(c) for c in rng.integers(0, n, dims))
        if dims == 2:
            loc = s.global_to_real(*coords)
            assert s.real_to_global(*loc) == coords and loc == o.global_to_real(*coords)
            kloc = s.global_to_reciprocal(*coords)
            assert s.reciprocal_to_global(*kloc) == coords and kloc == o.global_to_reciprocal(*coords)
            assert all(0 <= c < p for c in loc[:2]) and all(0 <= c < l for c in loc[2:])
        else:
            loc = s.global_to_real_3d(*coords)
            assert s.real_to_global_3d(*loc) == coords and loc == o.global_to_real_3d(*coords)
            kloc = s.global_to_reciprocal_3d(*coords)
            assert s.reciprocal_to_global_3d(*kloc) == coords and kloc == o.global_to_reciprocal_3d(*coords)
    for c in rng.integers(0, l, 10):
        assert s.block_cyclic_to_packed(s.packed_to_block_cyclic(int(c))) == int(c)
        assert s.lp_inv(s.lp(int(c))) == int(c)
    for q in rng.integers(0, n, 10):
        assert s.ct_inv(s.ct(int(q))) == int(q)


@settings(**dict(SETTINGS, max_examples=40))
@given(shapes(max_log2n=(7, 4)))
def test_oracle_forward_is_numpy_fftn(oracle_mod, shape):
    """The oracle's block-cyclic forward transform is the natural-layout DFT,
    read through the reciprocal map; its inverse brings the field back."""
    dims, n, p, b = shape
    s = oracle_mod.Shape(dims, n, p, b)
    nat = uniform_field((n,) * dims, n + 7 * p + b)
    y = s.forward_fft(s.natural_to_real(nat))
    assert rel_max(s.reciprocal_to_natural(y), np.fft.fftn(nat)) < 1e-12
    back = s.real_to_natural(s.inverse_fft(y, 1.0 / n ** dims))
    assert rel_max(back, nat) < 1e-12
