"""Device-resident API: plans, gridding, degridding and the dirty image on
torch tensors that already live on the GPU.

PyTorch is used for device memory, streams and the FFT only; every other
operation is a libavgrid.so kernel.  All calls are asynchronous on the
current torch stream.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import torch

from . import _lib
from ._lib import (AV_DEGRID_F32_PRODUCTS, AV_DEGRID_INCREMENTAL, AV_F32, AV_F64, AV_GEOM_CONJUGATE, AV_GEOM_DEGRID,
                   AV_PLAN_ROW_CHANNELS,
                   AV_GEOM_REFERENCE, AV_GRID_ACCUMULATE, AV_GRID_ADJOINT,
                   AV_GRID_COMPENSATED, AV_GRID_F32_PRODUCTS, AV_GRID_PSF, AV_GRID_SUBTILE,
                   call)


def require_gpu(device=None) -> torch.device:
    """The GPU device to run on; raises when no GPU is visible."""
    if not torch.cuda.is_available():
        raise RuntimeError("astroviper_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    _lib.lib()
    if device is None:
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device(device)


def _ptr(t: Optional[torch.Tensor]):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


def _aligned(t: Optional[torch.Tensor], nbytes: int) -> Optional[torch.Tensor]:
    """`t`, or a copy of it when its data does not start on an `nbytes`
    boundary: the kernels read a sample's polarisations with vector loads
    (16 bytes of weights or visibilities, 4 bytes of flags)."""
    if t is None or t.data_ptr() % nbytes == 0:
        return t
    return t.clone()


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _real_code(dtype: torch.dtype) -> int:
    if dtype in (torch.float32, torch.complex64):
        return AV_F32
    if dtype in (torch.float64, torch.complex128):
        return AV_F64
    raise TypeError(f"unsupported dtype {dtype}")


def real_dtype(code: int) -> torch.dtype:
    return torch.float32 if code == AV_F32 else torch.float64


def complex_dtype(code: int) -> torch.dtype:
    return torch.complex64 if code == AV_F32 else torch.complex128


# ------------------------------------------------------------ PS tables --
def ps_function(u: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """prolate_spheroidal_function on a float64 device tensor."""
    u = u.contiguous().to(torch.float64)
    g = torch.empty_like(u)
    k = torch.empty_like(u)
    call("av_ps_function", _ptr(u), u.numel(), _ptr(g), _ptr(k), _stream())
    return g, k


def ps_kernel_1d(oversampling: int, support: int, device=None) -> torch.Tensor:
    """create_prolate_spheroidal_kernel_1D as a float64 device tensor."""
    dev = require_gpu(device)
    out = torch.empty(oversampling * (support // 2 + 1), dtype=torch.float64, device=dev)
    call("av_ps_kernel_1d", int(oversampling), int(support), _ptr(out), _stream())
    return out


def ps_correction_1d(n: int, device=None) -> torch.Tensor:
    """grdsf(|2 * coordinates(n)|): one axis of the gridding-correction taper."""
    dev = require_gpu(device)
    out = torch.empty(n, dtype=torch.float64, device=dev)
    call("av_ps_correction_1d", int(n), _ptr(out), _stream())
    return out


def ps_outer(k1: torch.Tensor, nfill: int) -> torch.Tensor:
    """4-D kernel of create_prolate_spheroidal_kernel from its 1-D rows."""
    k1 = k1.contiguous().to(torch.float64)
    nrow, support = k1.shape
    out = torch.empty((nrow, nrow, support, support), dtype=torch.float64, device=k1.device)
    call("av_ps_outer", _ptr(k1), nrow, int(nfill), support, _ptr(out), _stream())
    return out


# ----------------------------------------------------------------- plans --
class GridPlan:
    """Binning of one set of uvw coordinates into (uv tile, image channel)
    bins on the GPU; shared by :meth:`grid` and :meth:`degrid`.

    Parameters mirror the reference gridder's: ``n_uv`` is the grid shape
    (axis 2, axis 3), ``cell`` the signed pixel increment in radians,
    ``support`` / ``oversampling`` the kernel table.  ``geometry`` is
    ``"reference"`` (gridding and degridding each use their reference
    rounding rule), ``"degrid"`` (both use the degridder's, making
    ``grid(adjoint=True)`` the exact adjoint of ``degrid``) or
    ``"conjugate"`` (the reference grid rule with every sample placed at its
    conjugate centre and no degrid records: the second scatter of
    ``do_imaging_weight``, standard_grid.py:389-430).  With
    ``row_channels`` every row is one sample at its own channel (``row_chan``
    in :meth:`set_uvw`), and the plan holds n_rows samples instead of
    n_rows * n_chan.  ``n_baseline`` > 1 says the rows are time-major with
    that many baselines per time step (``av_plan_set_row_order``): the
    binning then writes its records baseline by baseline.
    """

    def __init__(self, n_rows: int, n_chan: int, n_img_chan: int, n_uv, cell,
                 support: int = 7, oversampling: int = 100, tile=(16, 16),
                 geometry: str = "reference", unit_cap: int = 0, device=None,
                 row_channels: bool = False, n_baseline: int = 1):
        self.device = require_gpu(device)
        self.n_rows, self.n_chan, self.n_img_chan = int(n_rows), int(n_chan), int(n_img_chan)
        self.n_u, self.n_v = int(n_uv[0]), int(n_uv[1])
        self.cell = (float(cell[0]), float(cell[1]))
        self.support, self.oversampling = int(support), int(oversampling)
        self.tile = (int(tile[0]), int(tile[1]))
        geo = {"reference": AV_GEOM_REFERENCE, "degrid": AV_GEOM_DEGRID,
               "conjugate": AV_GEOM_REFERENCE | AV_GEOM_CONJUGATE}[geometry]
        if row_channels:
            geo |= AV_PLAN_ROW_CHANNELS
        self.geometry = geometry
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            call("av_plan_create", self.n_rows, self.n_chan, self.n_img_chan, self.n_u,
                 self.n_v, self.cell[0], self.cell[1], self.support, self.oversampling,
                 self.tile[0], self.tile[1], geo, int(unit_cap), ctypes.byref(h))
        self._h = h
        if int(n_baseline) > 1:
            call("av_plan_set_row_order", self._h, int(n_baseline))
        self._keep = ()
        self.row_chan = None
        self.cgk = ps_kernel_1d(self.oversampling, self.support, self.device)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().av_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_uvw(self, uvw: torch.Tensor, freq: torch.Tensor,
                chan_map: Optional[torch.Tensor] = None,
                row_chan: Optional[torch.Tensor] = None) -> "GridPlan":
        """Bin uvw [n_rows, 3] float64 at frequencies freq [n_chan] float64.

        chan_map [n_chan] int32 maps visibility to image channels (default
        identity for cube, zeros for continuum).  With row_chan [n_rows]
        int32 each row is one sample at freq[row_chan[r]].
        The plan keeps references to these tensors until the next call.
        """
        uvw = uvw.to(self.device, torch.float64).contiguous()
        freq = freq.to(self.device, torch.float64).contiguous()
        if uvw.shape != (self.n_rows, 3):
            raise ValueError(f"uvw must be [{self.n_rows}, 3], got {tuple(uvw.shape)}")
        if freq.numel() != self.n_chan:
            raise ValueError("freq must have n_chan entries")
        if chan_map is not None:
            chan_map = chan_map.to(self.device, torch.int32).contiguous()
        if row_chan is not None:
            row_chan = row_chan.to(self.device, torch.int32).contiguous()
        self._keep = (uvw, freq, chan_map, row_chan)
        self.row_chan = row_chan
        call("av_plan_set_uvw", self._h, _ptr(uvw), _ptr(freq), _ptr(chan_map),
             _ptr(row_chan), _stream())
        return self

    @property
    def n_samples(self) -> int:
        return self.n_rows if self.row_chan is not None else self.n_rows * self.n_chan

    def stats(self) -> dict:
        out = (ctypes.c_int64 * 8)()
        call("av_plan_stats", self._h, ctypes.cast(out, ctypes.c_void_p), _stream())
        keys = ("samples", "grid_valid", "degrid_valid", "records", "bins",
                "grid_units", "degrid_units", "zero_list")
        return dict(zip(keys, list(out)))

    # -------------------------------------------------------------------
    def grid(self, vis: Optional[torch.Tensor], weight: Optional[torch.Tensor],
             flag: Optional[torch.Tensor] = None, grid: Optional[torch.Tensor] = None,
             sumwt: Optional[torch.Tensor] = None, *, n_pol: Optional[int] = None,
             psf: bool = False, accumulate: bool = False, compensated: bool = False,
             adjoint: bool = False, dtype: Optional[torch.dtype] = None,
             subtile: bool = False, f32_products: bool = False):
        """Grid samples onto [n_img_chan, n_pol, n_u, n_v] complex planes.

        vis [samples, n_pol] complex, weight [samples, n_pol] real, flag
        [samples, n_pol] uint8/bool or None.  Returns (grid, sumwt).
        FP32 16 x 16 tiles run the split-FP16 matrix-core kernel (FP16
        halves of every operand, FP32 sums) unless ``f32_products`` asks for
        the FP32 16x16x4 kernel or ``subtile`` for the sub-tile scheduled
        4x4x1 one.
        """
        ref = vis if vis is not None else weight
        if ref is None:
            raise ValueError("need vis or weight")
        code = _real_code(dtype if dtype is not None else ref.dtype)
        if n_pol is None:
            n_pol = ref.shape[-1]
        ns = self.n_samples
        if vis is not None and not psf:
            vis = vis.reshape(ns, n_pol)
            if vis.dtype != complex_dtype(code):
                raise TypeError(f"vis must be {complex_dtype(code)}")
        if weight is not None and not adjoint:
            weight = weight.reshape(ns, n_pol)
            if weight.dtype != real_dtype(code):
                raise TypeError(f"weight must be {real_dtype(code)}")
        if flag is not None:
            flag = flag.reshape(ns, n_pol)
            if flag.dtype == torch.bool:
                flag = flag.view(torch.uint8)
        vis = _aligned(vis, min(16, 2 * n_pol * real_dtype(code).itemsize))
        weight = _aligned(weight, min(16, n_pol * real_dtype(code).itemsize))
        flag = _aligned(flag, n_pol)
        shape = (self.n_img_chan, n_pol, self.n_u, self.n_v)
        if grid is None:
            grid = torch.empty(shape, dtype=complex_dtype(code), device=self.device)
            if accumulate:
                grid.zero_()
        elif tuple(grid.shape) != shape or grid.dtype != complex_dtype(code):
            raise ValueError(f"grid must be {shape} {complex_dtype(code)}")
        if sumwt is None:
            sumwt = torch.zeros((self.n_img_chan, n_pol), dtype=torch.float64,
                                device=self.device)
        if ns == 0:  # nothing to grid: an empty (zeroed) result
            if not accumulate:
                grid.zero_()
                sumwt.zero_()
            return grid, sumwt
        mode = ((AV_GRID_PSF if psf else 0) | (AV_GRID_ACCUMULATE if accumulate else 0)
                | (AV_GRID_COMPENSATED if compensated else 0)
                | (AV_GRID_ADJOINT if adjoint else 0) | (AV_GRID_SUBTILE if subtile else 0)
                | (AV_GRID_F32_PRODUCTS if f32_products else 0))
        call("av_grid", self._h, _ptr(None if psf else vis),
             _ptr(None if adjoint else weight), _ptr(flag), int(n_pol), _ptr(self.cgk),
             _ptr(grid), _ptr(sumwt), code, mode, _stream())
        return grid, sumwt

    def degrid(self, grid: torch.Tensor, flag: Optional[torch.Tensor] = None,
               vis_out: Optional[torch.Tensor] = None, *, n_pol: Optional[int] = None,
               pol_map: Optional[torch.Tensor] = None, incremental: bool = False,
               f32_products: bool = False):
        """Predict samples [samples, n_pol] from grid [n_img_chan, npol_g, n_u, n_v].
        FP32 16 x 16 tiles run the split-FP16 matrix-core degridder unless
        ``f32_products`` asks for FP32 products."""
        code = _real_code(grid.dtype)
        npol_g = grid.shape[1]
        if n_pol is None:
            n_pol = npol_g if vis_out is None else vis_out.shape[-1]
        ns = self.n_samples
        if vis_out is None:
            vis_out = torch.empty((ns, n_pol), dtype=grid.dtype, device=self.device)
            if incremental:
                vis_out.zero_()
        if flag is not None:
            flag = flag.reshape(ns, n_pol)
            if flag.dtype == torch.bool:
                flag = flag.view(torch.uint8)
        flag = _aligned(flag, n_pol)
        if pol_map is not None:
            pol_map = pol_map.to(self.device, torch.int32).contiguous()
        if ns == 0:
            return vis_out
        mode = ((AV_DEGRID_INCREMENTAL if incremental else 0)
                | (AV_DEGRID_F32_PRODUCTS if f32_products else 0))
        call("av_degrid", self._h, _ptr(grid.contiguous()), int(npol_g), _ptr(pol_map),
             _ptr(flag), int(n_pol), _ptr(self.cgk), _ptr(vis_out), code, mode, _stream())
        return vis_out


# ------------------------------------------------------ kernel choice --
def grid_kernel_name(tile, n_pol: int, fp64: bool = False, compensated: bool = False,
                     support: int = 7, oversampling: int = 100,
                     lds_bytes: int = 160 * 1024, subtile: bool = False,
                     f32_products: bool = False) -> str:
    """Name of the tile kernel av_grid runs for this plan shape: for FP32
    16 x 16 tiles without compensated sums, the sub-tile scheduled kernel
    (4x4x1 products) when ``subtile`` is asked for, else the split-FP16
    kernel unless ``f32_products`` is asked for (both when their LDS
    fits), else the dense
    matrix-core kernel for tiles of 16 x 16 multiples whose result blocks
    ((tile_u/16) * 2*n_pol * (tile_v/16)) fit its register budget (16 blocks
    for FP32, 8 for FP64 and for compensated sums) and whose LDS (a weight
    table padded by 16 * max(tile_u, tile_v)/16 zero rows per side, the FP64
    kernel-sum table and four wavefronts' staging) fits in a workgroup's
    `lds_bytes`, else the LDS kernel."""
    tu, tv = int(tile[0]), int(tile[1])
    if (tu, tv) == (16, 16) and not fp64 and not compensated and subtile:
        nw = 2 * (oversampling // 2 + 2) + 1
        up16 = lambda b: (b + 15) & ~15
        wave = up16(up16(65 * 8 * n_pol) + 65 * 4) + 16 * 64 + 16
        total = up16(max(support + 8, 16) * nw * 4) + up16(nw * 8) + 4 * wave
        if total <= lds_bytes:
            return "k_grid_sub"
    limit = 8 if (fp64 or compensated) else 16
    if tu in (16, 32) and tv in (16, 32, 64) and (tu // 16) * 2 * n_pol * (tv // 16) <= limit:
        r = 8 if fp64 else 4
        nw = 2 * (oversampling // 2 + 2) + 1
        pad = max(tu, tv)
        up16 = lambda b: (b + 15) & ~15
        total = up16((support + 2 * pad) * nw * r) + up16(nw * 8) + 4 * (64 * 2 * n_pol * r + 64 * 8)
        if total <= lds_bytes:
            if (tu, tv) == (16, 16) and not fp64 and not compensated and not f32_products:
                return "k_grid_h3"
            return "k_grid_mfma"
    return "k_grid_tiles"


def degrid_kernel_name(tile, n_pol: int, support: int = 7, fp64: bool = False,
                       f32_products: bool = False) -> str:
    """Name of the kernel av_degrid runs: for FP32 16 x 16 tiles whose
    rows plus the halo of support//2 + 1 pixels fit 24 region rows the
    split-FP16 matrix-core kernel (unless ``f32_products``), else the FP32 or
    FP64 matrix-core kernel when the tile plus its halo spans at most 16
    blocks of 16 components per row, else the LDS kernel."""
    halo = support // 2 + 1
    blocks = ((int(tile[1]) + 2 * halo) * 2 * n_pol + 15) // 16
    if (not fp64 and not f32_products and (int(tile[0]), int(tile[1])) == (16, 16)
            and support >= 2 and 16 + 2 * halo <= 24 and blocks <= 12):
        return "k_degrid_h3"
    return "k_degrid_mfma" if blocks <= 16 else "k_degrid_tiles"


# --------------------------------------------------------- dirty image --
def dirty_image(grid: torch.Tensor, sumwt: torch.Tensor, support: int = 7,
                oversampling: int = 100, out_dtype: torch.dtype = torch.float64,
                fft_dtype: torch.dtype = torch.complex128,
                work: Optional[torch.Tensor] = None) -> torch.Tensor:
    """real(fftshift(ifft2(ifftshift(grid)))) / taper * nx*ny / sumwt per plane
    (standard_grid.py:693-708).  grid [..., n_u, n_v] complex on the GPU.

    The FFT runs in `fft_dtype` (FP64 by default, also for an FP32 grid: the
    taper correction divides by down to grdsf(1)^2 = 1.6e-5 at the image
    edge and would magnify FP32 FFT rounding to ~1e-3 of the peak).

    FP64 transforms of even-sized planes go through a complex-to-real FFT:
    only the real part of the inverse FFT is kept, and it is the transform of
    the grid's Hermitian half spectrum, which av_ifftshift_hermitian folds
    (with both shifts) in one pass; the FFT then does half the work and the
    passes around it move half the bytes.  Other cases transform the whole
    complex planes (`work`, if given, holds them).
    """
    code = _real_code(grid.dtype)
    fcode = _real_code(fft_dtype)
    n_u, n_v = grid.shape[-2:]
    planes = grid.numel() // (n_u * n_v)
    g = grid.contiguous()
    tu = ps_correction_1d(n_u, grid.device)
    tv = tu if n_v == n_u else ps_correction_1d(n_v, grid.device)
    sw = sumwt.contiguous().to(torch.float64)
    if fft_dtype == torch.complex128 and n_u % 2 == 0 and n_v % 2 == 0:
        half = torch.empty(g.shape[:-1] + (n_v // 2 + 1,), dtype=torch.complex128,
                           device=g.device)
        call("av_ifftshift_hermitian", _ptr(g), _ptr(half), planes, n_u, n_v, code, _stream())
        img = torch.fft.irfft2(half, s=(n_u, n_v), dim=(-2, -1))
        del half
        out = img if out_dtype == torch.float64 else torch.empty(grid.shape, dtype=out_dtype,
                                                                 device=grid.device)
        call("av_normalise_real", _ptr(img), _ptr(sw), _ptr(tu), _ptr(tv), planes, n_u, n_v,
             _ptr(out), _real_code(out_dtype), _stream())
        return out
    if work is None:
        work = torch.empty(g.shape, dtype=fft_dtype, device=g.device)
    call("av_ifftshift", _ptr(g), _ptr(work), planes, n_u, n_v, code, fcode, _stream())
    torch.fft.ifft2(work, dim=(-2, -1), out=work)
    out = torch.empty(grid.shape, dtype=out_dtype, device=grid.device)
    call("av_normalise", _ptr(work), _ptr(sw), _ptr(tu), _ptr(tv), planes, n_u, n_v, fcode,
         _ptr(out), _real_code(out_dtype), _stream())
    return out


# ------------------------------------------------------ plain baseline --
def grid_plain(uvw: torch.Tensor, freq: torch.Tensor, vis, weight, flag, n_uv, cell,
               n_img_chan: int, chan_map, support: int = 7, oversampling: int = 100,
               grid=None, sumwt=None, psf: bool = False, cgk=None, chans=None):
    """Plain one-thread-per-sample gridder, launched once per channel.
    vis/weight/flag [rows, n_chan, n_pol]."""
    code = _real_code(weight.dtype)
    n_rows, n_chan, n_pol = weight.shape
    dev = weight.device
    if cgk is None:
        cgk = ps_kernel_1d(oversampling, support, dev)
    if grid is None:
        grid = torch.zeros((n_img_chan, n_pol, int(n_uv[0]), int(n_uv[1])),
                           dtype=complex_dtype(code), device=dev)
    if sumwt is None:
        sumwt = torch.zeros((n_img_chan, n_pol), dtype=torch.float64, device=dev)
    fl = flag.view(torch.uint8) if (flag is not None and flag.dtype == torch.bool) else flag
    freq_h = freq.detach().cpu().tolist()
    for c in (range(n_chan) if chans is None else chans):
        call("av_grid_plain", _ptr(uvw), n_rows, n_chan, c, float(freq_h[c]),
             int(chan_map[c]), _ptr(None if psf else vis), _ptr(weight), _ptr(fl), n_pol,
             _ptr(cgk), support, oversampling, float(cell[0]), float(cell[1]),
             int(n_uv[0]), int(n_uv[1]), _ptr(grid), _ptr(sumwt), code,
             AV_GRID_PSF if psf else 0, _stream())
    return grid, sumwt


def degrid_plain(uvw: torch.Tensor, freq: torch.Tensor, grid, flag, cell, chan_map,
                 n_pol: int, support: int = 7, oversampling: int = 100, vis_out=None,
                 cgk=None, chans=None):
    """Plain one-thread-per-sample degridder, launched once per channel."""
    code = _real_code(grid.dtype)
    n_rows = uvw.shape[0]
    n_chan = freq.numel()
    dev = grid.device
    if cgk is None:
        cgk = ps_kernel_1d(oversampling, support, dev)
    if vis_out is None:
        vis_out = torch.zeros((n_rows, n_chan, n_pol), dtype=grid.dtype, device=dev)
    fl = flag.view(torch.uint8) if (flag is not None and flag.dtype == torch.bool) else flag
    freq_h = freq.detach().cpu().tolist()
    n_u, n_v = grid.shape[-2:]
    for c in (range(n_chan) if chans is None else chans):
        call("av_degrid_plain", _ptr(uvw), n_rows, n_chan, c, float(freq_h[c]),
             int(chan_map[c]), _ptr(grid), grid.shape[1], _ptr(fl), n_pol, _ptr(cgk),
             support, oversampling, float(cell[0]), float(cell[1]), n_u, n_v,
             _ptr(vis_out), code, 0, _stream())
    return vis_out
