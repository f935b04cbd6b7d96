"""ctypes binding of libavgrid.so (the C ABI declared in include/avgrid.h).

The library is required: nothing in this package computes without it, and
every entry point raises if it is missing or if no GPU is visible.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libavgrid.so")

AV_ABI_VERSION = 12  # must equal AV_ABI_VERSION of include/avgrid.h

AV_OK = 0
AV_ERR_INVALID = -1
AV_ERR_HIP = -2
AV_ERR_NOMEM = -3
AV_ERR_UNSUPPORTED = -4

AV_F32 = 1
AV_F64 = 2

AV_GRID_PSF = 1
AV_GRID_ACCUMULATE = 2
AV_GRID_COMPENSATED = 4
AV_GRID_ADJOINT = 8
AV_GRID_SUBTILE = 16
AV_GRID_F32_PRODUCTS = 32
AV_DEGRID_INCREMENTAL = 1
AV_DEGRID_F32_PRODUCTS = 2

AV_GEOM_REFERENCE = 0
AV_GEOM_DEGRID = 1
AV_PLAN_ROW_CHANNELS = 16
AV_GEOM_CONJUGATE = 32

_c = ctypes
_P = _c.c_void_p
_I32 = _c.c_int32
_I64 = _c.c_int64
_D = _c.c_double

# name -> (restype, argtypes)
SIGNATURES = {
    "av_last_error": (_c.c_char_p, []),
    "av_version": (_c.c_int, []),
    "av_device_info": (_c.c_int, [_P, _P, _P]),
    "av_ps_function": (_c.c_int, [_P, _I64, _P, _P, _P]),
    "av_ps_kernel_1d": (_c.c_int, [_I32, _I32, _P, _P]),
    "av_ps_correction_1d": (_c.c_int, [_I32, _P, _P]),
    "av_ps_taper": (_c.c_int, [_I32, _I32, _P, _P]),
    "av_ps_outer": (_c.c_int, [_P, _I32, _I32, _I32, _P, _P]),
    "av_plan_create": (_c.c_int, [_I64, _I32, _I32, _I32, _I32, _D, _D, _I32, _I32,
                                  _I32, _I32, _I32, _I32, _c.POINTER(_P)]),
    "av_plan_destroy": (_c.c_int, [_P]),
    "av_plan_set_uvw": (_c.c_int, [_P, _P, _P, _P, _P, _P]),
    "av_plan_set_degrid_transform": (_c.c_int, [_P, _D, _D, _D, _D]),
    "av_plan_set_row_order": (_c.c_int, [_P, _I32]),
    "av_plan_stats": (_c.c_int, [_P, _P, _P]),
    "av_grid": (_c.c_int, [_P, _P, _P, _P, _I32, _P, _P, _P, _I32, _I32, _P]),
    "av_degrid": (_c.c_int, [_P, _P, _I32, _P, _P, _I32, _P, _P, _I32, _I32, _P]),
    "av_grid_noalloc": (_c.c_int, [_P, _I64, _P, _I32, _P, _I32, _I32, _I32, _D, _D, _I32, _I32,
                                   _I32, _I32, _P, _P, _P, _I32, _P, _P, _P, _I32, _I32, _P]),
    "av_mosaic_grid": (_c.c_int, [_P, _P, _P, _I32, _I32, _P, _P, _P, _P, _I32, _I32, _I32,
                                  _I32, _I32, _P, _I32, _I32, _P, _P, _I32, _I32, _P]),
    "av_grid_plain": (_c.c_int, [_P, _I64, _I32, _I32, _D, _I32, _P, _P, _P, _I32, _P,
                                 _I32, _I32, _D, _D, _I32, _I32, _P, _P, _I32, _I32, _P]),
    "av_degrid_plain": (_c.c_int, [_P, _I64, _I32, _I32, _D, _I32, _P, _I32, _P, _I32,
                                   _P, _I32, _I32, _D, _D, _I32, _I32, _P, _I32, _I32,
                                   _P]),
    "av_ifftshift": (_c.c_int, [_P, _P, _I64, _I32, _I32, _I32, _I32, _P]),
    "av_shift_taper": (_c.c_int, [_P, _P, _I64, _I32, _I32, _I32, _I32, _P, _P, _I32, _D,
                                  _I32, _I32, _I32, _I32, _P]),
    "av_normalise": (_c.c_int, [_P, _P, _P, _P, _I64, _I32, _I32, _I32, _P, _I32, _P]),
    "av_ifftshift_hermitian": (_c.c_int, [_P, _P, _I64, _I32, _I32, _I32, _P]),
    "av_normalise_real": (_c.c_int, [_P, _P, _P, _P, _I64, _I32, _I32, _P, _I32, _P]),
    "av_hermitian_expand": (_c.c_int, [_P, _P, _I64, _I32, _I32, _I32, _I32, _P]),
    "av_hogbom_work_bytes": (_c.c_int, [_P]),
    "av_hogbom_maximg": (_c.c_int, [_P, _P, _I32, _I32, _I32, _P, _P]),
    "av_hogbom_clean": (_c.c_int, [_P, _P, _P, _P, _I32, _I32, _I32, _I32, _I32, _I32, _I32,
                                   _I32, _I32, _I32, _I32, _D, _D, _D, _I32, _I32, _I32, _P,
                                   _P]),
    "av_grid_imaging_weights": (_c.c_int, [_P, _I64, _I32, _P, _I32, _P, _I32, _I32, _I32, _D,
                                           _D, _P, _P, _P]),
    "av_degrid_imaging_weights": (_c.c_int, [_P, _P, _P, _I64, _I32, _P, _I32, _P, _I32, _I32,
                                             _D, _D, _P, _P, _P]),
    "av_gcf_aperture_patterns": (_c.c_int, [_P, _I32, _P, _P, _I32, _I32, _I32, _I32, _I32, _D,
                                            _D, _I32, _I32, _P, _P]),
    "av_gcf_phase_gradient": (_c.c_int, [_P, _I32, _I32, _I32, _I32, _I32, _P, _P]),
}

_lib = None


class AvgridError(RuntimeError):
    """A HIP runtime failure reported by libavgrid.so."""


def lib():
    """Load libavgrid.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m astroviper_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        ver = L.av_version()
        if ver != AV_ABI_VERSION:
            raise RuntimeError(f"{LIB_PATH} has ABI version {ver}, the bindings need "
                               f"{AV_ABI_VERSION}: rebuild with `python -m astroviper_amd.build`")
        _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    if rc == AV_OK:
        return
    msg = lib().av_last_error().decode(errors="replace")
    text = f"{what}: {msg}"
    if rc == AV_ERR_INVALID:
        raise ValueError(text)
    if rc == AV_ERR_UNSUPPORTED:
        raise NotImplementedError(text)
    if rc == AV_ERR_NOMEM:
        raise MemoryError(text)
    raise AvgridError(text)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)
