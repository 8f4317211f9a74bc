"""Python host interface over the C ABI: shapes, index maps and plans.

Names follow the reference's zmodel API (zmodel/utils.h, zmodel/transforms/fft.h,
zmodel/stencils/laplace.h).  Device buffers are torch tensors on the GPU (torch
is only the allocator and stream provider); every compute call goes to
libstencilfft.so.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

from . import _lib as L

PRECISIONS = {"fp64": L.ZM_FP64, "fp32": L.ZM_FP32, "fp32_mfma": L.ZM_FP32_MFMA}


@dataclass(frozen=True)
class Shape:
    """(dims, N, P, B) of the interleaved block-cyclic distribution."""

    dims: int
    n: int
    p: int
    b: int

    def _c(self) -> L.ZmShape:
        return L.ZmShape(self.dims, self.n, self.p, self.b)

    def check(self) -> "Dims":
        d = L.ZmDims()
        L.check(L._lib().zm_shape_check(ctypes.byref(self._c()), ctypes.byref(d)), f"shape {self}")
        return Dims(d.l, d.vec, d.vp, d.m, d.nranks, d.local, d.field)

    # ---- host index maps (zmodel/utils.h:122-320) ----
    def packed_to_block_cyclic(self, c: int) -> int:
        return L._lib().zm_packed_to_block_cyclic(ctypes.byref(self._c()), c)

    def block_cyclic_to_packed(self, c: int) -> int:
        return L._lib().zm_block_cyclic_to_packed(ctypes.byref(self._c()), c)

    def lp(self, i: int) -> int:
        return L._lib().zm_lp(ctypes.byref(self._c()), i)

    def ct(self, p: int) -> int:
        return L._lib().zm_ct(ctypes.byref(self._c()), p)

    def lp_inv(self, j: int) -> int:
        return L._lib().zm_lp_inv(ctypes.byref(self._c()), j)

    def ct_inv(self, k: int) -> int:
        return L._lib().zm_ct_inv(ctypes.byref(self._c()), k)

    def _map(self, name, ins, nout):
        outs = [ctypes.c_int() for _ in range(nout)]
        getattr(L._lib(), name)(ctypes.byref(self._c()), *ins, *[ctypes.byref(o) for o in outs])
        return tuple(o.value for o in outs)

    def real_to_global(self, rid, cid, lr, lc):
        return self._map("zm_real_to_global", (rid, cid, lr, lc), 2)

    def global_to_real(self, gr, gc):
        return self._map("zm_global_to_real", (gr, gc), 4)

    def real_to_global_3d(self, rid, cid, did, ld, lr, lc):
        return self._map("zm_real_to_global_3d", (rid, cid, did, ld, lr, lc), 3)

    def global_to_real_3d(self, gd, gr, gc):
        return self._map("zm_global_to_real_3d", (gd, gr, gc), 6)

    def reciprocal_to_global(self, rid, cid, lr, lc):
        return self._map("zm_reciprocal_to_global", (rid, cid, lr, lc), 2)

    def global_to_reciprocal(self, ky, kx):
        return self._map("zm_global_to_reciprocal", (ky, kx), 4)

    def reciprocal_to_global_3d(self, rid, cid, did, ld, lr, lc):
        return self._map("zm_reciprocal_to_global_3d", (rid, cid, did, ld, lr, lc), 3)

    def global_to_reciprocal_3d(self, kd, kr, kc):
        return self._map("zm_global_to_reciprocal_3d", (kd, kr, kc), 6)


@dataclass(frozen=True)
class Dims:
    l: int
    vec: int
    vp: int
    m: int
    nranks: int
    local: int
    field: int


def _ptr(t) -> int:
    if t is None:
        return None
    return t.data_ptr()


def _stream(stream) -> int:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class Plan:
    """A transform / stencil plan on the current GPU (zm_plan_create).

    ``precision`` is "fp64" (the reference's), "fp32" or "fp32_mfma".
    Every method takes torch tensors holding ``batch`` contiguous fields of
    N^d elements (complex128 for fp64, complex64 otherwise; float64 / float32
    for real stencil inputs) and returns the output tensor.
    """

    def __init__(self, dims: int, n: int, p: int, b: int, precision: str = "fp64", batch: int = 1,
                 two_stage: bool = False, caller_scratch: bool = False):
        self.shape = Shape(dims, n, p, b)
        self.dims = self.shape.check()
        self.precision = precision
        self.batch = batch
        self.caller_scratch = caller_scratch
        flags = (L.ZM_FLAG_TWO_STAGE if two_stage else 0) | (L.ZM_FLAG_CALLER_SCRATCH if caller_scratch else 0)
        h = ctypes.c_void_p()
        L.check(L._lib().zm_plan_create(ctypes.byref(h), ctypes.byref(self.shape._c()), PRECISIONS[precision],
                                       batch, flags), f"zm_plan_create({self.shape}, {precision})")
        self._h = h
        self.scratch_bytes = L._lib().zm_scratch_bytes(h)
        path = ctypes.c_uint()
        L.check(L._lib().zm_plan_path(h, ctypes.byref(path)), "zm_plan_path")
        self.two_stage = bool(path.value & L.ZM_PATH_TWO_STAGE)
        self.fast = bool(path.value & L.ZM_PATH_FAST_LINE)
        self.mfma = bool(path.value & L.ZM_PATH_MFMA)

    # ---- lifetime ----
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.check(L._lib().zm_plan_destroy(self._h), "zm_plan_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    # ---- tensors ----
    @property
    def complex_dtype(self):
        import torch
        return torch.complex128 if self.precision == "fp64" else torch.complex64

    @property
    def real_dtype(self):
        import torch
        return torch.float64 if self.precision == "fp64" else torch.float32

    def numel(self) -> int:
        return self.batch * self.dims.field

    def empty(self, real: bool = False, device="cuda"):
        import torch
        return torch.empty(self.numel(), dtype=self.real_dtype if real else self.complex_dtype, device=device)

    def _check_tensor(self, t, real=False):
        want = self.real_dtype if real else self.complex_dtype
        if t.dtype != want or t.numel() != self.numel() or not t.is_contiguous() or not t.is_cuda:
            raise ValueError(f"expected a contiguous CUDA tensor of {self.numel()} {want}, got "
                             f"{tuple(t.shape)} {t.dtype} on {t.device}")

    def _out(self, x, out, real=False):
        self._check_tensor(x, real)
        if out is None:
            import torch
            out = torch.empty_like(x)
        else:
            self._check_tensor(out, real)
        return out

    @staticmethod
    def _span(t):
        start = t.data_ptr()
        return start, start + t.numel() * t.element_size()

    def _scratch(self, scratch, *tensors):
        """Pointer of a caller's scratch tensor after checking it: a contiguous
        CUDA tensor on the input's device, at least scratch_bytes long, that
        overlaps none of `tensors`.  None selects the plan's own scratch (an
        error for a caller-scratch plan).  Raises before anything is launched."""
        if scratch is None:
            if self.caller_scratch:
                raise ValueError("this plan owns no scratch (caller_scratch=True): pass scratch=")
            return None
        nbytes = scratch.numel() * scratch.element_size()
        if not scratch.is_cuda or not scratch.is_contiguous() or nbytes < self.scratch_bytes:
            raise ValueError(f"scratch must be a contiguous CUDA tensor of at least {self.scratch_bytes} bytes, got "
                             f"{tuple(scratch.shape)} {scratch.dtype} ({nbytes} bytes) on {scratch.device}")
        if tensors and scratch.device != tensors[0].device:
            raise ValueError(f"scratch is on {scratch.device}, the fields on {tensors[0].device}")
        s0, s1 = self._span(scratch)
        for t in tensors:
            t0, t1 = self._span(t)
            if s0 < t1 and t0 < s1:
                raise ValueError("scratch overlaps an input or output field")
        return scratch.data_ptr()

    # ---- transforms ----
    def forward(self, x, out=None, scratch=None, stream=None):
        """forward_fft (zmodel/transforms/fft.cu:199-339 / fft3d.cu:234-412)."""
        out = self._out(x, out)
        sp = self._scratch(scratch, x, out)
        L.check(L._lib().zm_forward(self._h, _ptr(x), _ptr(out), sp, _stream(stream)), "zm_forward")
        return out

    def inverse(self, x, scale: float = 1.0, out=None, scratch=None, stream=None):
        """inverse_fft (zmodel/transforms/fft.cu:409-514 / fft3d.cu:508-633)."""
        out = self._out(x, out)
        sp = self._scratch(scratch, x, out)
        L.check(L._lib().zm_inverse(self._h, _ptr(x), _ptr(out), float(scale), sp, _stream(stream)),
                "zm_inverse")
        return out

    def laplace_forward(self, x, scale: float = 1.0, out=None, scratch=None, stream=None):
        """Forward transform of scale * Laplacian(x): the Laplacian + forward_fft
        of zmodel/main.cu:50-63, bit for bit forward(laplace(x, scale)); the
        Laplacian is formed inside the x pass where the plan has that pass."""
        out = self._out(x, out)
        sp = self._scratch(scratch, x, out)
        L.check(L._lib().zm_laplace_forward(self._h, _ptr(x), _ptr(out), float(scale), sp, _stream(stream)),
                "zm_laplace_forward")
        return out

    def permute_forward(self, x, out=None, scratch=None, stream=None):
        """forward_fft<false>: the data movement alone."""
        out = self._out(x, out)
        sp = self._scratch(scratch, x, out)
        L.check(L._lib().zm_permute_forward(self._h, _ptr(x), _ptr(out), sp, _stream(stream)),
                "zm_permute_forward")
        return out

    def permute_inverse(self, x, out=None, scratch=None, stream=None):
        """inverse_fft<false>: the data movement alone."""
        out = self._out(x, out)
        sp = self._scratch(scratch, x, out)
        L.check(L._lib().zm_permute_inverse(self._h, _ptr(x), _ptr(out), sp, _stream(stream)),
                "zm_permute_inverse")
        return out

    def forward_adjoint(self, x, out=None, scratch=None, stream=None):
        out = self._out(x, out)
        sp = self._scratch(scratch, x, out)
        L.check(L._lib().zm_forward_adjoint(self._h, _ptr(x), _ptr(out), sp, _stream(stream)),
                "zm_forward_adjoint")
        return out

    def inverse_adjoint(self, x, scale: float = 1.0, out=None, scratch=None, stream=None):
        out = self._out(x, out)
        sp = self._scratch(scratch, x, out)
        L.check(L._lib().zm_inverse_adjoint(self._h, _ptr(x), _ptr(out), float(scale), sp, _stream(stream)),
                "zm_inverse_adjoint")
        return out

    # ---- stencils ----
    def _elem(self, x):
        return L.ZM_COMPLEX if x.is_complex() else L.ZM_REAL

    def laplace(self, x, scale: float = 1.0, out=None, stream=None):
        """5-point (2D) / 7-point (3D) periodic Laplacian, times `scale`."""
        out = self._out(x, out, real=not x.is_complex())
        L.check(L._lib().zm_laplace(self._h, _ptr(x), _ptr(out), float(scale), self._elem(x), _stream(stream)),
                "zm_laplace")
        return out

    def dx4(self, x, delta: float = 1.0, out=None, stream=None, adjoint: bool = False):
        out = self._out(x, out, real=not x.is_complex())
        fn = L._lib().zm_dx4_adjoint if adjoint else L._lib().zm_dx4
        L.check(fn(self._h, _ptr(x), _ptr(out), float(delta), self._elem(x), _stream(stream)), "zm_dx4")
        return out

    def dy4(self, x, delta: float = 1.0, out=None, stream=None, adjoint: bool = False):
        out = self._out(x, out, real=not x.is_complex())
        fn = L._lib().zm_dy4_adjoint if adjoint else L._lib().zm_dy4
        L.check(fn(self._h, _ptr(x), _ptr(out), float(delta), self._elem(x), _stream(stream)), "zm_dy4")
        return out

    def dz4(self, x, delta: float = 1.0, out=None, stream=None, adjoint: bool = False):
        """4th-order z (depth) derivative; 3D plans only."""
        out = self._out(x, out, real=not x.is_complex())
        fn = L._lib().zm_dz4_adjoint if adjoint else L._lib().zm_dz4
        L.check(fn(self._h, _ptr(x), _ptr(out), float(delta), self._elem(x), _stream(stream)), "zm_dz4")
        return out

    SPECTRAL_KINDS = {"poisson": L.ZM_SPECTRAL_POISSON, "heat": L.ZM_SPECTRAL_HEAT, "scale": L.ZM_SPECTRAL_SCALE,
                      "stencil": L.ZM_SPECTRAL_STENCIL, "poisson_stencil": L.ZM_SPECTRAL_POISSON_STENCIL}

    def spectral_multiply(self, x, kind: str, p0: float = 0.0, p1: float = 0.0, length: float = 1.0, out=None,
                          stream=None, adjoint: bool = False):
        """out = m(k) x on the reciprocal layout.  "stencil" / "poisson_stencil"
        take p0 = Laplacian scale and p1 = dx4 delta (0: no dx4 term)."""
        if out is None:
            out = self._out(x, None)
        else:
            self._check_tensor(out)
        fn = L._lib().zm_spectral_multiply_adjoint if adjoint else L._lib().zm_spectral_multiply
        L.check(fn(self._h, _ptr(x), _ptr(out), self.SPECTRAL_KINDS[kind], float(p0), float(p1), float(length),
                   _stream(stream)), "zm_spectral_multiply")
        return out

    # ---- per-stage timing (zm_timing_*) ----
    def timing(self, enable: bool = True):
        L.check(L._lib().zm_timing_enable(self._h, 1 if enable else 0), "zm_timing_enable")

    def timing_read(self) -> dict:
        """{stage name: (launch count, total ms)} for every stage with launches."""
        out = {}
        for k, name in enumerate(L.STAGES):
            c = ctypes.c_int()
            ms = ctypes.c_double()
            L.check(L._lib().zm_timing_read(self._h, k, ctypes.byref(c), ctypes.byref(ms)), "zm_timing_read")
            if c.value:
                out[name] = (c.value, ms.value)
        return out

    def timing_reset(self):
        L.check(L._lib().zm_timing_reset(self._h), "zm_timing_reset")

    # ---- device index maps and layout conversion ----
    def real_index_table(self, stream=None):
        import torch
        t = torch.empty(self.dims.field, dtype=torch.int64, device="cuda")
        L.check(L._lib().zm_real_index_table(self._h, _ptr(t), _stream(stream)), "zm_real_index_table")
        return t

    def reciprocal_index_table(self, stream=None):
        import torch
        t = torch.empty(self.dims.field, dtype=torch.int64, device="cuda")
        L.check(L._lib().zm_reciprocal_index_table(self._h, _ptr(t), _stream(stream)), "zm_reciprocal_index_table")
        return t

    def natural_to_packed(self, x, out=None, stream=None):
        out = self._out(x, out)
        L.check(L._lib().zm_natural_to_packed(self._h, _ptr(x), _ptr(out), _stream(stream)), "zm_natural_to_packed")
        return out

    def packed_to_natural(self, x, out=None, stream=None):
        out = self._out(x, out)
        L.check(L._lib().zm_packed_to_natural(self._h, _ptr(x), _ptr(out), _stream(stream)), "zm_packed_to_natural")
        return out
