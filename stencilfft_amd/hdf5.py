"""HDF5 field I/O through libstencilfft_hdf5.so (include/stencilfft_hdf5.h).

Mirrors store_host_hdf5 / init_host_hdf5 and their _3d forms
(zmodel/utils.cu:103-281) for the whole virtual grid: one N^d dataset of the
compound type {"real": double, "imag": double} in global natural order, each
rank's part the block-cyclic hyperslab of its coordinates, vector-packed in
memory.

Host buffers are complex128 numpy arrays of P^d * L^d values in world-rank
order (the layout of a plan's real-space field).  A torch tensor is accepted
for storing (it is copied to host memory first).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._lib import HERE, ZmShape, check

HDF5_LIB_PATH = os.path.join(HERE, "libstencilfft_hdf5.so")

_SP = ctypes.POINTER(ZmShape)
_DP = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "zm_store_hdf5": (ctypes.c_int, [_SP, _DP, ctypes.c_char_p, ctypes.c_char_p]),
    "zm_load_hdf5": (ctypes.c_int, [_SP, _DP, ctypes.c_char_p, ctypes.c_char_p]),
}

_handle = None


def _lib() -> ctypes.CDLL:
    """Load libstencilfft_hdf5.so (raises if it has not been built)."""
    global _handle
    if _handle is None:
        from ._lib import _lib as core
        core()  # resolves libstencilfft.so (zm_shape_check, zm_strerror) first
        if not os.path.exists(HDF5_LIB_PATH):
            raise ImportError(
                f"{HDF5_LIB_PATH} is missing: build it with `python -m stencilfft_amd.build` "
                "on a machine with the HDF5 C library (set HDF5_HOME)")
        h = ctypes.CDLL(HDF5_LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _handle = h
    return _handle


def _shape(shape) -> ZmShape:
    """A Shape, a Plan (its shape) or a (dims, n, p, b) tuple."""
    from .plan import Plan, Shape
    if isinstance(shape, Plan):
        shape = shape.shape
    if isinstance(shape, Shape):
        return shape._c()
    dims, n, p, b = shape
    return ZmShape(dims, n, p, b)


def _expected_size(s: ZmShape) -> int:
    return s.n ** s.dims


def store_host_hdf5(shape, grid, filename: str, dataset_name: str = "data") -> None:
    """Write the packed virtual grid `grid` (P^d * L^d complex values) to
    `filename` (created or truncated) as dataset `dataset_name`."""
    s = _shape(shape)
    if hasattr(grid, "detach"):  # torch tensor
        grid = grid.detach().to("cpu").numpy()
    a = np.ascontiguousarray(np.asarray(grid), dtype=np.complex128).reshape(-1)
    if a.size != _expected_size(s):
        raise ValueError(f"grid has {a.size} values, the shape needs {_expected_size(s)}")
    check(_lib().zm_store_hdf5(ctypes.byref(s), a.ctypes.data_as(_DP), os.fsencode(filename),
                               dataset_name.encode()), "zm_store_hdf5")


def init_host_hdf5(shape, filename: str, dataset_name: str = "data") -> np.ndarray:
    """Read dataset `dataset_name` of `filename` (N^d compound values) into a new
    packed virtual grid (complex128, P^d * L^d values)."""
    s = _shape(shape)
    out = np.empty(_expected_size(s), dtype=np.complex128)
    check(_lib().zm_load_hdf5(ctypes.byref(s), out.ctypes.data_as(_DP), os.fsencode(filename),
                              dataset_name.encode()), "zm_load_hdf5")
    return out
