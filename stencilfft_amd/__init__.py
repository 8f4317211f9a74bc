"""MI355X-native interleaved block-cyclic FFT and stencils (StencilFFT zmodel path).

The compute lives in ``libstencilfft.so`` (hand-written gfx950 HIP kernels behind
the C ABI of ``include/stencilfft.h``); this package is the Python host layer.
"""
from ._lib import LIB_PATH, ZmError  # noqa: F401
from .plan import Dims, Plan, Shape  # noqa: F401

__all__ = ["Plan", "Shape", "Dims", "ZmError", "LIB_PATH", "version"]


def version() -> str:
    from ._lib import _lib
    return _lib().zm_version().decode()
