"""CPU FP64 oracle for StencilFFT's zmodel hot path -- TEST INFRASTRUCTURE ONLY.

This package is the oracle.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import it; the product package
(``stencilfft_amd``) never does.  It wraps ``oracle/liboracle.so`` (built from
``oracle/oracle.c`` by ``oracle/Makefile``), a plain-C restatement of the
reference's stages; each C function cites the reference file:line it follows.

Grids are numpy arrays holding the whole virtual grid: the P^d rank-local
buffers concatenated in world-rank order.  Complex grids are ``complex128``
arrays of ``N^d`` elements; real grids ``float64``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "liboracle.so"], check=True)
    return _LIB_PATH


class _Shape(ctypes.Structure):
    _fields_ = [
        ("dims", ctypes.c_int), ("n", ctypes.c_int), ("p", ctypes.c_int), ("b", ctypes.c_int),
        ("l", ctypes.c_int), ("vec", ctypes.c_int), ("vp", ctypes.c_int), ("m", ctypes.c_int),
        ("nranks", ctypes.c_int), ("local", ctypes.c_int64), ("total", ctypes.c_int64),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        sp = ctypes.POINTER(_Shape)
        L = _lib
        L.zo_shape_init.argtypes = [sp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        for name in ("zo_packed_to_block_cyclic", "zo_block_cyclic_to_packed", "zo_lp", "zo_ct", "zo_lp_inv", "zo_ct_inv"):
            getattr(L, name).argtypes = [sp, ctypes.c_int]
            getattr(L, name).restype = ctypes.c_int
        L.zo_real_to_global.argtypes = [sp] + [ctypes.c_int] * 4 + [ip] * 2
        L.zo_global_to_real.argtypes = [sp] + [ctypes.c_int] * 2 + [ip] * 4
        L.zo_reciprocal_to_global.argtypes = [sp] + [ctypes.c_int] * 4 + [ip] * 2
        L.zo_global_to_reciprocal.argtypes = [sp] + [ctypes.c_int] * 2 + [ip] * 4
        L.zo_real_to_global_3d.argtypes = [sp] + [ctypes.c_int] * 6 + [ip] * 3
        L.zo_global_to_real_3d.argtypes = [sp] + [ctypes.c_int] * 3 + [ip] * 6
        L.zo_reciprocal_to_global_3d.argtypes = [sp] + [ctypes.c_int] * 6 + [ip] * 3
        L.zo_global_to_reciprocal_3d.argtypes = [sp] + [ctypes.c_int] * 3 + [ip] * 6
        for name in ("zo_natural_to_real", "zo_real_to_natural", "zo_reciprocal_to_natural", "zo_natural_to_reciprocal"):
            getattr(L, name).argtypes = [sp, dp, dp, ctypes.c_int]
        L.zo_vector_pack.argtypes = [sp, dp, dp, ctypes.c_int64, ctypes.c_int]
        L.zo_vector_unpack.argtypes = [sp, dp, dp, ctypes.c_int64, ctypes.c_int]
        L.zo_init_host.argtypes = [sp, dp, ctypes.c_int]
        L.zo_twiddle_table.argtypes = [sp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.zo_set_dft_mode.argtypes = [ctypes.c_int]
        L.zo_get_dft_mode.restype = ctypes.c_int
        L.zo_dft_many.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int]
        for name in ("zo_fft1a", "zo_fft1b"):
            getattr(L, name).argtypes = [sp, dp, dp, ctypes.c_int]
        L.zo_pack_forward.argtypes = [sp, dp, dp]
        L.zo_pack_inverse.argtypes = [sp, dp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.zo_alltoall.argtypes = [sp, ctypes.c_int, dp, dp]
        L.zo_unpack_forward.argtypes = [sp, ctypes.c_int, dp, dp, ctypes.c_int]
        L.zo_unpack_inverse.argtypes = [sp, ctypes.c_int, dp, dp, ctypes.c_int]
        for name in ("zo_repack_transpose", "zo_inverse_repack_transpose", "zo_repack_transpose_xy_3d",
                     "zo_repack_transpose_yz_3d", "zo_inverse_repack_transpose_xy_3d",
                     "zo_inverse_repack_transpose_yz_3d"):
            getattr(L, name).argtypes = [sp, dp, dp]
        L.zo_forward_fft.argtypes = [sp, dp, dp, ctypes.c_int, ctypes.POINTER(dp)]
        L.zo_inverse_fft.argtypes = [sp, dp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.POINTER(dp)]
        L.zo_laplace.argtypes = [sp, dp, dp, ctypes.c_int]
        L.zo_dx.argtypes = [sp, dp, dp, ctypes.c_double, ctypes.c_int]
        L.zo_dx.restype = ctypes.c_int
        L.zo_dy.argtypes = [sp, dp, dp, ctypes.c_double, ctypes.c_int]
        L.zo_dy.restype = ctypes.c_int
    return _lib


def _dp(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ncomp(a: np.ndarray) -> int:
    if a.dtype == np.complex128:
        return 2
    if a.dtype == np.float64:
        return 1
    raise TypeError(f"oracle works in FP64 only, got {a.dtype}")


class Shape:
    """Validated (dims, N, P, B) with the reference's derived macros."""

    def __init__(self, dims: int, n: int, p: int, b: int):
        self._s = _Shape()
        if lib().zo_shape_init(ctypes.byref(self._s), dims, n, p, b) != 0:
            raise ValueError(f"invalid shape dims={dims} N={n} P={p} B={b}")
        s = self._s
        self.dims, self.n, self.p, self.b = s.dims, s.n, s.p, s.b
        self.l, self.vec, self.vp, self.m = s.l, s.vec, s.vp, s.m
        self.nranks, self.local, self.total = s.nranks, s.local, s.total

    @property
    def ref(self):
        return ctypes.byref(self._s)

    def __repr__(self):
        return f"Shape(dims={self.dims}, N={self.n}, P={self.p}, B={self.b})"

    # ---------------- scalar index maps ----------------
    def _call(self, name, ins, nout):
        outs = [ctypes.c_int() for _ in range(nout)]
        getattr(lib(), name)(self.ref, *ins, *[ctypes.byref(o) for o in outs])
        return tuple(o.value for o in outs)

    def packed_to_block_cyclic(self, c):
        return lib().zo_packed_to_block_cyclic(self.ref, c)

    def block_cyclic_to_packed(self, c):
        return lib().zo_block_cyclic_to_packed(self.ref, c)

    def lp(self, i):
        return lib().zo_lp(self.ref, i)

    def ct(self, p):
        return lib().zo_ct(self.ref, p)

    def lp_inv(self, j):
        return lib().zo_lp_inv(self.ref, j)

    def ct_inv(self, k):
        return lib().zo_ct_inv(self.ref, k)

    def real_to_global(self, rid, cid, lr, lc):
        return self._call("zo_real_to_global", (rid, cid, lr, lc), 2)

    def global_to_real(self, gr, gc):
        return self._call("zo_global_to_real", (gr, gc), 4)

    def reciprocal_to_global(self, rid, cid, lr, lc):
        return self._call("zo_reciprocal_to_global", (rid, cid, lr, lc), 2)

    def global_to_reciprocal(self, ky, kx):
        return self._call("zo_global_to_reciprocal", (ky, kx), 4)

    def real_to_global_3d(self, rid, cid, did, ld, lr, lc):
        return self._call("zo_real_to_global_3d", (rid, cid, did, ld, lr, lc), 3)

    def global_to_real_3d(self, gd, gr, gc):
        return self._call("zo_global_to_real_3d", (gd, gr, gc), 6)

    def reciprocal_to_global_3d(self, rid, cid, did, ld, lr, lc):
        return self._call("zo_reciprocal_to_global_3d", (rid, cid, did, ld, lr, lc), 3)

    def global_to_reciprocal_3d(self, kd, kr, kc):
        return self._call("zo_global_to_reciprocal_3d", (kd, kr, kc), 6)

    # ---------------- whole-grid conversions ----------------
    def natural_shape(self):
        return (self.n,) * self.dims

    def _conv(self, name, src, out_shape):
        src = np.ascontiguousarray(src)
        out = np.empty(out_shape, dtype=src.dtype)
        getattr(lib(), name)(self.ref, _dp(src), _dp(out), _ncomp(src))
        return out

    def natural_to_real(self, nat):
        return self._conv("zo_natural_to_real", nat, (self.total,))

    def real_to_natural(self, grid):
        return self._conv("zo_real_to_natural", grid, self.natural_shape())

    def reciprocal_to_natural(self, grid):
        return self._conv("zo_reciprocal_to_natural", grid, self.natural_shape())

    def natural_to_reciprocal(self, nat):
        return self._conv("zo_natural_to_reciprocal", nat, (self.total,))

    def init_host(self, dtype=np.complex128):
        out = np.zeros(self.total, dtype=dtype)
        lib().zo_init_host(self.ref, _dp(out), _ncomp(out))
        return out

    def twiddle_table(self, rank_coord: int, direction: int = -1, index_mode: bool = False):
        out = np.zeros(self.l * self.l, dtype=np.complex128)
        lib().zo_twiddle_table(self.ref, rank_coord, direction, 1 if index_mode else 0, _dp(out))
        return out

    # ---------------- stages ----------------
    def fft1a(self, x, direction=-1):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        out = np.empty_like(x)
        lib().zo_fft1a(self.ref, _dp(x), _dp(out), direction)
        return out

    def fft1b(self, x, direction=-1):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        out = np.empty_like(x)
        lib().zo_fft1b(self.ref, _dp(x), _dp(out), direction)
        return out

    def pack_forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        out = np.empty_like(x)
        lib().zo_pack_forward(self.ref, _dp(x), _dp(out))
        return out

    def pack_inverse(self, x, scale=1.0 + 0j, do_compute=True):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        out = np.empty_like(x)
        lib().zo_pack_inverse(self.ref, _dp(x), _dp(out), float(np.real(scale)), float(np.imag(scale)), int(do_compute))
        return out

    def alltoall(self, axis, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        out = np.empty_like(x)
        lib().zo_alltoall(self.ref, axis, _dp(x), _dp(out))
        return out

    def unpack_forward(self, axis, x, do_compute=True):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        out = np.empty_like(x)
        lib().zo_unpack_forward(self.ref, axis, _dp(x), _dp(out), int(do_compute))
        return out

    def unpack_inverse(self, axis, x, do_compute=True):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        out = np.empty_like(x)
        lib().zo_unpack_inverse(self.ref, axis, _dp(x), _dp(out), int(do_compute))
        return out

    def _perm(self, name, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        out = np.empty_like(x)
        getattr(lib(), name)(self.ref, _dp(x), _dp(out))
        return out

    def repack_transpose(self, x):
        return self._perm("zo_repack_transpose", x)

    def inverse_repack_transpose(self, x):
        return self._perm("zo_inverse_repack_transpose", x)

    def repack_transpose_xy_3d(self, x):
        return self._perm("zo_repack_transpose_xy_3d", x)

    def repack_transpose_yz_3d(self, x):
        return self._perm("zo_repack_transpose_yz_3d", x)

    def inverse_repack_transpose_xy_3d(self, x):
        return self._perm("zo_inverse_repack_transpose_xy_3d", x)

    def inverse_repack_transpose_yz_3d(self, x):
        return self._perm("zo_inverse_repack_transpose_yz_3d", x)

    def _nstages(self):
        return 11 if self.dims == 2 else 17

    def forward_fft(self, x, do_compute=True, capture_stages=False):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        out = np.empty_like(x)
        stages = None
        sptr = None
        if capture_stages:
            stages = [np.empty_like(x) for _ in range(self._nstages())]
            sptr = (ctypes.POINTER(ctypes.c_double) * len(stages))(*[_dp(a) for a in stages])
        lib().zo_forward_fft(self.ref, _dp(x), _dp(out), int(do_compute), sptr)
        return (out, stages) if capture_stages else out

    def inverse_fft(self, x, scale=1.0 + 0j, do_compute=True, capture_stages=False):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        out = np.empty_like(x)
        stages = None
        sptr = None
        if capture_stages:
            stages = [np.empty_like(x) for _ in range(self._nstages())]
            sptr = (ctypes.POINTER(ctypes.c_double) * len(stages))(*[_dp(a) for a in stages])
        lib().zo_inverse_fft(self.ref, _dp(x), _dp(out), float(np.real(scale)), float(np.imag(scale)),
                             int(do_compute), sptr)
        return (out, stages) if capture_stages else out

    # ---------------- stencils ----------------
    def laplace(self, x):
        x = np.ascontiguousarray(x)
        out = np.empty_like(x)
        lib().zo_laplace(self.ref, _dp(x), _dp(out), _ncomp(x))
        return out

    def dx(self, x, delta=1.0):
        x = np.ascontiguousarray(x)
        out = np.empty_like(x)
        if lib().zo_dx(self.ref, _dp(x), _dp(out), float(delta), _ncomp(x)) != 0:
            raise ValueError("dx needs a 2D shape with B >= 2")
        return out

    def dy(self, x, delta=1.0):
        x = np.ascontiguousarray(x)
        out = np.empty_like(x)
        if lib().zo_dy(self.ref, _dp(x), _dp(out), float(delta), _ncomp(x)) != 0:
            raise ValueError("dy needs a 2D shape with B >= 2")
        return out

    def deriv4(self, x, axis: str, delta=1.0):
        """dx / dy / dz of a packed grid (axis "x", "y" or "z"), 2D or 3D, by
        way of the natural layout and ``deriv4_natural``."""
        if self.b < 2 or (axis == "z" and self.dims != 3):
            raise ValueError("deriv4 needs B >= 2, and dz a 3D shape")
        nat_axis = {"x": self.dims - 1, "y": self.dims - 2, "z": 0}[axis]
        return self.natural_to_real(deriv4_natural(self.real_to_natural(x), nat_axis, delta))


def deriv4_natural(f: np.ndarray, axis: int, delta: float = 1.0) -> np.ndarray:
    """Periodic 4th-order central difference along `axis` of a natural-layout
    field, summed in the reference's order (zmodel/stencils/dx.cu:29-79,
    dy.cu:34-97): s = 0 + f(-1) * -8 + f(-2) + f(+1) * 8 - f(+2), then
    s / (12 delta), each operation rounded once.  Complex fields: the real
    stencil on re and im separately.  In 2D this is bit-identical to the C
    restatement of dx / dy; along depth (no reference counterpart) it is the
    same stencil on the third axis."""
    if np.iscomplexobj(f):
        out = np.empty_like(f)
        out.real = deriv4_natural(np.ascontiguousarray(f.real), axis, delta)
        out.imag = deriv4_natural(np.ascontiguousarray(f.imag), axis, delta)
        return out
    s = np.zeros_like(f)
    s = s + np.roll(f, 1, axis) * -8.0
    s = s + np.roll(f, 2, axis)
    s = s + np.roll(f, -1, axis) * 8.0
    s = s - np.roll(f, -2, axis)
    return s / (12.0 * delta)


def set_dft_mode(mode: int) -> None:
    """0 = direct DFT (definition, long-double accumulation), 1 = radix-2 FFT."""
    lib().zo_set_dft_mode(int(mode))


def get_dft_mode() -> int:
    return lib().zo_get_dft_mode()


def dft_many(x: np.ndarray, n: int, count: int, dist: int, stride: int, direction: int = -1) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(x)
    lib().zo_dft_many(_dp(x), _dp(out), n, count, dist, stride, direction)
    return out
