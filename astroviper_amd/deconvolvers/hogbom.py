"""Hogbom CLEAN on the GPU: mirror of core/imaging/deconvolvers/hogbom (the
_hogbom_ext extension built from src/hclean.cpp and python/bindings.cpp),
with the same functions, arguments, defaults and result dictionary.  The
iterations run in libavgrid.so's cooperative CLEAN kernel.

Precision follows the extension's overload resolution (bindings.cpp:243-301,
float32 overloads registered first): the float64 kernel runs when every
array passed is float64 and C-contiguous, the float32 kernel otherwise (the
arrays are then converted, as pybind11's forcecast does).
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np
import torch

from .. import _lib
from .. import device as _dev

__all__ = ["maximg", "clean", "clean_nc", "get_dtype_name", "is_float32", "is_float64",
           "clean_device"]

_work_bytes = None


def _work(dev) -> torch.Tensor:
    global _work_bytes
    if _work_bytes is None:
        n = ctypes.c_int64()
        _lib.call("av_hogbom_work_bytes", ctypes.byref(n))
        _work_bytes = int(n.value)
    return torch.empty(_work_bytes, dtype=torch.uint8, device=dev)


def _given(a) -> bool:
    return a is not None and np.asarray(a).size > 0


def _dtype_for(*arrays) -> type:
    """The overload pybind11 picks: float64 only when every given array
    already is a C-contiguous float64 array."""
    for a in arrays:
        if a is None:
            continue
        a = np.asarray(a)
        if a.size == 0 and a.dtype == np.float64:
            continue
        if a.dtype != np.float64 or not a.flags.c_contiguous:
            return np.float32
    return np.float64


def _code(dt) -> int:
    return _lib.AV_F32 if dt == np.float32 else _lib.AV_F64


def get_dtype_name(arr) -> str:
    """bindings.cpp:287-289: the array's dtype string."""
    return np.asarray(arr).dtype.str


def is_float32(arr) -> bool:
    return np.asarray(arr).dtype == np.float32


def is_float64(arr) -> bool:
    return np.asarray(arr).dtype == np.float64


def _device_maximg(img: torch.Tensor, mask: Optional[torch.Tensor], nx: int, ny: int):
    out = torch.empty(2, dtype=img.dtype, device=img.device)
    _lib.call("av_hogbom_maximg", _dev._ptr(img), _dev._ptr(mask), nx, ny,
              _code(np.float32 if img.dtype == torch.float32 else np.float64), _dev._ptr(out),
              _dev._stream())
    lo, hi = out.cpu().numpy()
    return lo, hi


def maximg(image, mask=None):
    """maximg (bindings.cpp:16-57 -> hclean.cpp:51-81): (min, max) of the
    image over mask > 0.5, in the image's precision."""
    img = np.asarray(image)
    dt = _dtype_for(img, mask)
    if img.ndim != 2:
        raise RuntimeError("Image must be 2D array")
    ny, nx = img.shape
    m = None
    if _given(mask):
        m = np.asarray(mask)
        if m.ndim != 2:
            raise RuntimeError("Mask must be 2D array")
        if m.shape != (ny, nx):
            raise RuntimeError("Mask dimensions must match image spatial dimensions")
    dev = _dev.require_gpu()
    ti = torch.from_numpy(np.ascontiguousarray(img, dtype=dt)).to(dev)
    tm = None if m is None else torch.from_numpy(np.ascontiguousarray(m, dtype=dt)).to(dev)
    lo, hi = _device_maximg(ti, tm, nx, ny)
    return dt(lo), dt(hi)


def _box(clean_box, nx, ny):
    """bindings.cpp:108-127: -1 means the full extent; clamped to the image."""
    xbeg, xend, ybeg, yend = 0, nx, 0, ny
    if len(clean_box) == 4:
        xbeg, xend, ybeg, yend = (int(v) for v in clean_box)
        xbeg = 0 if xbeg == -1 else xbeg
        xend = nx if xend == -1 else xend
        ybeg = 0 if ybeg == -1 else ybeg
        yend = ny if yend == -1 else yend
        xbeg = max(0, min(xbeg, nx - 1))
        xend = max(xbeg + 1, min(xend, nx))
        ybeg = max(0, min(ybeg, ny - 1))
        yend = max(ybeg + 1, min(yend, ny))
    return xbeg, xend, ybeg, yend


def clean_device(residual: torch.Tensor, psf: torch.Tensor, model: torch.Tensor,
                 mask: Optional[torch.Tensor] = None, box=None, max_iter: int = 100,
                 start_iter: int = 0, gain: float = 0.1, threshold: float = 0.0,
                 speedup: float = 0.0, progress_callback=None, stop_callback=None,
                 work: Optional[torch.Tensor] = None) -> int:
    """CLEAN one [ny, nx] plane in place on device tensors (float32 or
    float64, all of one dtype): residual -= components * psf, model +=
    components.  Returns the iteration count reached (hclean.cpp:106-210).
    The callbacks are called where the reference calls them: at the start
    iteration and every max(1, (max_iter - start_iter) // 10) iterations
    (iter % cycle == 1), then once at the end if any iteration ran."""
    ny, nx = residual.shape
    for t in (residual, psf, model) + ((mask,) if mask is not None else ()):
        if t.dtype != residual.dtype or tuple(t.shape) != (ny, nx) or not t.is_contiguous():
            raise ValueError("residual, psf, model and mask must be contiguous [ny, nx] "
                             "tensors of one dtype")
    dt = np.float32 if residual.dtype == torch.float32 else np.float64
    xbeg, xend, ybeg, yend = box if box is not None else (0, nx, 0, ny)
    if max_iter - start_iter > (1 << 32) // 1024 - 2:
        raise ValueError("too many iterations for one call")
    if work is None:
        work = _work(residual.device)
    report = bool(progress_callback is not None or stop_callback is not None)
    cycle = max(1, (max_iter - start_iter) // 10) if report else 0
    start, skip, px, py = start_iter, -1, 0, 0
    it, maxval = start_iter, 0.0
    while True:
        _lib.call("av_hogbom_clean", _dev._ptr(model), _dev._ptr(residual), _dev._ptr(psf),
                  _dev._ptr(mask), nx, ny, xbeg, xend, ybeg, yend, int(max_iter),
                  int(start_iter), int(start), cycle, int(skip), float(dt(gain)),
                  float(dt(threshold)), float(dt(speedup)), px, py, _code(dt), _dev._ptr(work),
                  _dev._stream())
        head = work[:32].cpu().numpy()
        it, px, py, reason = (int(x) for x in head[8:24].view(np.int32))
        maxval = dt(head[24:32].view(np.float64)[0])
        if reason < 0:
            raise _lib.AvgridError("av_hogbom_clean: a workgroup did not reach the grid barrier")
        if reason != 1:
            break
        if progress_callback is not None:
            try:
                progress_callback(it, px, py, maxval)
            except RuntimeError as e:  # bindings.cpp:143-149
                print("Warning: Progress callback failed:", e)
        stop = False
        if stop_callback is not None:
            try:
                stop = bool(stop_callback())
            except RuntimeError as e:  # bindings.cpp:155-162
                print("Warning: Stop callback failed:", e)
        if stop:
            break
        start, skip = it, it
    if it > start_iter and progress_callback is not None:
        try:
            progress_callback(it, px, py, maxval)
        except RuntimeError as e:
            print("Warning: Progress callback failed:", e)
    return it


def clean(dirty_image, psf, mask=None, clean_box=(-1, -1, -1, -1), max_iter: int = 100,
          start_iter: int = 0, gain: float = 0.1, threshold: float = 0.0, speedup: float = 0.0,
          progress_callback=None, stop_callback=None) -> dict:
    """clean (bindings.cpp:61-193): CLEAN a copy of dirty_image [ny, nx] with
    psf [ny, nx]; returns model_image, residual_image, iterations_performed,
    final_peak, total_flux_cleaned and converged."""
    d = np.asarray(dirty_image)
    p = np.asarray(psf)
    dt = _dtype_for(d, p, mask)
    if d.ndim != 2:
        raise RuntimeError("Dirty image must be 2D [ny, nx]")
    if p.ndim != 2:
        raise RuntimeError("PSF must be 2D [ny, nx]")
    ny, nx = d.shape
    if p.shape != (ny, nx):
        raise RuntimeError("PSF dimensions must match image spatial dimensions")
    m = None
    if _given(mask):
        m = np.asarray(mask)
        if m.ndim != 2 or m.shape != (ny, nx):
            raise RuntimeError("Mask must be 2D with same spatial dimensions as image")
    box = _box(clean_box, nx, ny)
    dev = _dev.require_gpu()
    resid = torch.from_numpy(np.ascontiguousarray(d, dtype=dt)).to(dev).clone()
    tp = torch.from_numpy(np.ascontiguousarray(p, dtype=dt)).to(dev)
    tm = None if m is None else torch.from_numpy(np.ascontiguousarray(m, dtype=dt)).to(dev)
    model = torch.zeros((ny, nx), dtype=resid.dtype, device=dev)
    it = clean_device(resid, tp, model, tm, box, max_iter, start_iter, gain, threshold, speedup,
                      progress_callback, stop_callback)
    lo, hi = _device_maximg(resid, tm, nx, ny)
    final_peak = dt(max(abs(dt(lo)), abs(dt(hi))))
    model_h = model.cpu().numpy()
    # the binding's sequential sum of |model| (bindings.cpp:175-178)
    total = np.cumsum(np.abs(model_h).ravel(), dtype=dt)[-1] if model_h.size else dt(0)
    return {"model_image": model_h, "residual_image": resid.cpu().numpy(),
            "iterations_performed": it, "final_peak": final_peak,
            "total_flux_cleaned": dt(total), "converged": bool(final_peak <= dt(threshold))}


def clean_nc(dirty_image, psf, gain: float = 0.1, threshold: float = 0.0, max_iter: int = 100,
             clean_box=(-1, -1, -1, -1)) -> dict:
    """clean_nc (bindings.cpp:199-216): clean without mask or callbacks."""
    return clean(dirty_image, psf, None, clean_box, max_iter, 0, gain, threshold, 0.0)
