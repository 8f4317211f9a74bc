"""ctypes binding of libstencilfft.so (the C ABI in include/stencilfft.h).

The library is loaded from this package directory.  There is no fallback: if
the shared library is missing, importing the compute entry points raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstencilfft.so")

ZM_OK = 0
(ZM_ERR_INVALID_ARG, ZM_ERR_SHAPE, ZM_ERR_UNSUPPORTED, ZM_ERR_DEVICE, ZM_ERR_ALIAS, ZM_ERR_LAUNCH,
 ZM_ERR_NOMEM, ZM_ERR_IO, ZM_ERR_SCRATCH_LIMIT) = range(1, 10)
ZM_NOALLOC_SCRATCH_BYTES = 64 << 20
STATUS_NAMES = {
    0: "ZM_OK",
    1: "ZM_ERR_INVALID_ARG",
    2: "ZM_ERR_SHAPE",
    3: "ZM_ERR_UNSUPPORTED",
    4: "ZM_ERR_DEVICE",
    5: "ZM_ERR_ALIAS",
    6: "ZM_ERR_LAUNCH",
    7: "ZM_ERR_NOMEM",
    8: "ZM_ERR_IO",
    9: "ZM_ERR_SCRATCH_LIMIT",
}

ZM_FP64, ZM_FP32, ZM_FP32_MFMA = 0, 1, 2
ZM_FLAG_TWO_STAGE = 1
ZM_FLAG_CALLER_SCRATCH = 2
ZM_PATH_TWO_STAGE, ZM_PATH_FAST_LINE, ZM_PATH_MFMA = 1, 2, 4
ZM_COMPLEX, ZM_REAL = 0, 1
ZM_SPECTRAL_POISSON, ZM_SPECTRAL_HEAT, ZM_SPECTRAL_SCALE, ZM_SPECTRAL_STENCIL, ZM_SPECTRAL_POISSON_STENCIL = range(5)
STAGES = ["line", "stage_a", "stage_b", "repack", "laplace", "deriv", "spectral", "convert", "line_laplace"]


class ZmShape(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int), ("n", ctypes.c_int), ("p", ctypes.c_int), ("b", ctypes.c_int)]


class ZmDims(ctypes.Structure):
    _fields_ = [
        ("l", ctypes.c_int), ("vec", ctypes.c_int), ("vp", ctypes.c_int), ("m", ctypes.c_int),
        ("nranks", ctypes.c_int), ("local", ctypes.c_int64), ("field", ctypes.c_int64),
    ]


class ZmError(RuntimeError):
    """A non-zero zm_status from the library."""

    def __init__(self, status: int, where: str = ""):
        self.status = status
        name = STATUS_NAMES.get(status, str(status))
        msg = _lib().zm_strerror(status).decode() if _handle is not None else name
        super().__init__(f"{where}: {name} ({msg})" if where else f"{name} ({msg})")


# Every exported symbol with (restype, argtypes).
_VP = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_IP = ctypes.POINTER(ctypes.c_int)
_SP = ctypes.POINTER(ZmShape)
SIGNATURES = {
    "zm_shape_check": (_I, [_SP, ctypes.POINTER(ZmDims)]),
    "zm_plan_create": (_I, [ctypes.POINTER(_VP), _SP, _I, _I, ctypes.c_uint]),
    "zm_plan_destroy": (_I, [_VP]),
    "zm_plan_dims": (_I, [_VP, _SP, ctypes.POINTER(ZmDims), _IP, _IP]),
    "zm_plan_path": (_I, [_VP, ctypes.POINTER(ctypes.c_uint)]),
    "zm_scratch_bytes": (ctypes.c_size_t, [_VP]),
    "zm_field_bytes": (ctypes.c_size_t, [_VP]),
    "zm_forward": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "zm_forward_noalloc": (_I, [_VP, _VP, _VP, _VP]),
    "zm_laplace_forward": (_I, [_VP, _VP, _VP, _D, _VP, _VP]),
    "zm_inverse": (_I, [_VP, _VP, _VP, _D, _VP, _VP]),
    "zm_permute_forward": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "zm_permute_inverse": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "zm_forward_adjoint": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "zm_inverse_adjoint": (_I, [_VP, _VP, _VP, _D, _VP, _VP]),
    "zm_laplace": (_I, [_VP, _VP, _VP, _D, _I, _VP]),
    "zm_dx4": (_I, [_VP, _VP, _VP, _D, _I, _VP]),
    "zm_dy4": (_I, [_VP, _VP, _VP, _D, _I, _VP]),
    "zm_dx4_adjoint": (_I, [_VP, _VP, _VP, _D, _I, _VP]),
    "zm_dy4_adjoint": (_I, [_VP, _VP, _VP, _D, _I, _VP]),
    "zm_dz4": (_I, [_VP, _VP, _VP, _D, _I, _VP]),
    "zm_dz4_adjoint": (_I, [_VP, _VP, _VP, _D, _I, _VP]),
    "zm_spectral_multiply": (_I, [_VP, _VP, _VP, _I, _D, _D, _D, _VP]),
    "zm_spectral_multiply_adjoint": (_I, [_VP, _VP, _VP, _I, _D, _D, _D, _VP]),
    "zm_packed_to_block_cyclic": (_I, [_SP, _I]),
    "zm_block_cyclic_to_packed": (_I, [_SP, _I]),
    "zm_real_to_global": (None, [_SP, _I, _I, _I, _I, _IP, _IP]),
    "zm_global_to_real": (None, [_SP, _I, _I, _IP, _IP, _IP, _IP]),
    "zm_real_to_global_3d": (None, [_SP, _I, _I, _I, _I, _I, _I, _IP, _IP, _IP]),
    "zm_global_to_real_3d": (None, [_SP, _I, _I, _I, _IP, _IP, _IP, _IP, _IP, _IP]),
    "zm_lp": (_I, [_SP, _I]),
    "zm_ct": (_I, [_SP, _I]),
    "zm_lp_inv": (_I, [_SP, _I]),
    "zm_ct_inv": (_I, [_SP, _I]),
    "zm_reciprocal_to_global": (None, [_SP, _I, _I, _I, _I, _IP, _IP]),
    "zm_global_to_reciprocal": (None, [_SP, _I, _I, _IP, _IP, _IP, _IP]),
    "zm_reciprocal_to_global_3d": (None, [_SP, _I, _I, _I, _I, _I, _I, _IP, _IP, _IP]),
    "zm_global_to_reciprocal_3d": (None, [_SP, _I, _I, _I, _IP, _IP, _IP, _IP, _IP, _IP]),
    "zm_init_host": (_I, [_SP, ctypes.POINTER(ctypes.c_double), _I]),
    "zm_real_index_table": (_I, [_VP, _VP, _VP]),
    "zm_reciprocal_index_table": (_I, [_VP, _VP, _VP]),
    "zm_natural_to_packed": (_I, [_VP, _VP, _VP, _VP]),
    "zm_packed_to_natural": (_I, [_VP, _VP, _VP, _VP]),
    "zm_timing_enable": (_I, [_VP, _I]),
    "zm_timing_read": (_I, [_VP, _I, _IP, ctypes.POINTER(ctypes.c_double)]),
    "zm_timing_reset": (_I, [_VP]),
    "zm_strerror": (ctypes.c_char_p, [_I]),
    "zm_version": (ctypes.c_char_p, []),
}

_handle = None


def _lib() -> ctypes.CDLL:
    """Load libstencilfft.so (raises if it has not been built)."""
    global _handle
    if _handle is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m stencilfft_amd.build` "
                "(there is no CPU fallback)")
        h = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _handle = h
    return _handle


def check(status: int, where: str = "") -> None:
    if status != ZM_OK:
        raise ZmError(status, where)
