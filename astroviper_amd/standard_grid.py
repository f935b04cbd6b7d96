"""Standard convolutional gridder on the GPU.

Mirror of core/imaging/imaging_utils/standard_grid.py: the same entry points,
arguments, defaults, in-place semantics and errors, with the gridding done by
libavgrid.so.  Host NumPy arrays go in and come out; the grid, the sums of
weights and the dirty image are computed on the GPU.

Precision follows the grid: complex128 grids (the reference default) are
accumulated in FP64, complex64 grids in FP32 (FP64 coordinates and sums of
weights either way).
"""
from __future__ import annotations

from typing import List, Tuple, Union

import numpy as np
import torch

from . import device as _dev
from .msv4 import datasets, is_datatree
from .staging import StagedGridder, array_partition

# Samples per staged block (rows x channels): bounds the device buffers and
# the plan of one call.
_BLOCK_SAMPLES = 1 << 24
# Plans are limited to 2^32/5 samples; larger degrid calls are cut into row
# blocks.
_MAX_PLAN_SAMPLES = 1 << 28


def _prec(dtype) -> int:
    dtype = np.dtype(dtype)
    if dtype in (np.dtype(np.complex64), np.dtype(np.float32)):
        return _dev.AV_F32
    return _dev.AV_F64


def _cgk_device(cgk_1D, oversampling, support, dev):
    need = oversampling * (support // 2 + 1)
    cgk = np.ascontiguousarray(np.asarray(cgk_1D, dtype=np.float64))
    if cgk.size < need:
        raise ValueError(f"cgk_1D has {cgk.size} entries, the kernel needs {need}")
    return torch.from_numpy(cgk[:need].copy()).to(dev)


def grid_to_device(vis_data, uvw, weight, flag, freq_chan, chan_map, n_img_chan, n_uv,
                   delta_lm, support, oversampling, do_psf, prec, cgk_dev,
                   grid_dev=None, sumwt_dev=None, tile=(16, 16), geometry="reference"):
    """Grid host arrays [nt, nb, nchan, npol] on the GPU, accumulating into
    grid_dev [n_img_chan, npol, n_u, n_v] / sumwt_dev when given.  The rows
    stream to the device in blocks through pinned buffers, with one plan
    reused for every block (staging.StagedGridder).
    ``geometry="conjugate"`` places every sample at its conjugate centre."""
    weight = np.asarray(weight)
    nt, nb, nchan, npol = weight.shape
    sg = StagedGridder(n_uv, delta_lm, nchan, npol, support=support, oversampling=oversampling,
                       fp32=prec == _dev.AV_F32, block_rows=max(nb, _BLOCK_SAMPLES // max(1, nchan)),
                       tile=tile, psf=do_psf, cgk=cgk_dev, geometry=geometry,
                       chan_map=chan_map, n_img_chan=n_img_chan, grid=grid_dev,
                       sumwt=sumwt_dev)
    try:
        sg.add(array_partition(None if do_psf else vis_data, np.asarray(uvw, dtype=np.float64),
                               weight, flag, np.asarray(freq_chan, dtype=np.float64)))
    finally:
        sg.close()
    return sg.grid, sg.sumwt


def standard_grid_jit(grid, sum_weight, do_psf, do_imaging_weight, vis_data, uvw, freq_chan,
                      chan_map, pol_map, weight, cgk_1D, n_uv, delta_lm, support, oversampling):
    """standard_grid.py:246-432 on the GPU: accumulate into `grid`
    (n_chan, n_pol, n_u, n_v) and `sum_weight` (n_chan, n_pol) in place.

    With ``do_imaging_weight`` every sample is gridded twice
    (standard_grid.py:389-430): at its centre and at its conjugate centre
    int(-u + N/2 + 0.5) with the same kernel weights (a second plan with the
    conjugate geometry); conjugate cells outside the grid are dropped, and
    the sum of weights counts both.  For the PSF with two or more
    polarisations every polarisation then grids (w0 + w1) / 2
    (standard_grid.py:356-360)."""
    weight = np.asarray(weight)
    npol = len(pol_map)
    if not np.array_equal(np.asarray(pol_map), np.arange(npol)) or weight.shape[3] != npol:
        raise NotImplementedError("pol_map must be arange(weight.shape[3])")
    if not np.iscomplexobj(grid) and not do_psf:
        raise TypeError("a real grid can only hold weights (do_psf=True)")
    support, oversampling = int(support), int(oversampling)
    n_img_chan = grid.shape[0]
    if grid.shape[1] != npol:
        raise ValueError("grid polarisation axis must match weight.shape[3]")
    prec = _prec(grid.dtype)
    dev = _dev.require_gpu()
    cgk = _cgk_device(cgk_1D, oversampling, support, dev)
    cdt = torch.complex64 if prec == _dev.AV_F32 else torch.complex128
    g_dev = torch.from_numpy(np.ascontiguousarray(grid)).to(dev).to(cdt)
    s_dev = torch.from_numpy(np.ascontiguousarray(sum_weight, dtype=np.float64)).to(dev)
    if do_imaging_weight and do_psf and npol >= 2:
        w2 = (weight[..., 0] + weight[..., 1]) / 2.0
        weight = np.ascontiguousarray(np.broadcast_to(w2[..., None], weight.shape))
    geometries = ("reference", "conjugate") if do_imaging_weight else ("reference",)
    for geo in geometries:
        g_dev, s_dev = grid_to_device(vis_data, uvw, weight, None, freq_chan, chan_map,
                                      n_img_chan, n_uv, delta_lm, support, oversampling, do_psf,
                                      prec, cgk, g_dev, s_dev, geometry=geo)
    g_h = g_dev.cpu().numpy()
    grid[...] = g_h if np.iscomplexobj(grid) else g_h.real
    sum_weight[...] = s_dev.cpu().numpy()


def standard_grid_numpy_wrap(vis_data, grid, sum_weight, uvw, weight, freq_chan, cgk_1D,
                             grid_params):
    """standard_grid.py:88-165: build the channel/polarisation maps, allocate
    the grid and sum of weights when both are empty, grid, return them."""
    n_chan = weight.shape[2]
    if grid_params["chan_mode"] == "cube":
        n_imag_chan = n_chan
        chan_map = (np.arange(0, n_chan)).astype(int)
    else:  # continuum
        n_imag_chan = 1
        chan_map = (np.zeros(n_chan)).astype(int)
    n_imag_pol = weight.shape[3]
    pol_map = (np.arange(0, n_imag_pol)).astype(int)
    n_uv = grid_params["image_size_padded"]
    delta_lm = grid_params["cell_size"]
    oversampling = grid_params["oversampling"]
    support = grid_params["support"]
    if grid.size == 0 and sum_weight.size == 0:
        if grid_params["complex_grid"]:
            grid = np.zeros((n_imag_chan, n_imag_pol, n_uv[0], n_uv[1]), dtype=np.complex128)
        else:
            grid = np.zeros((n_imag_chan, n_imag_pol, n_uv[0], n_uv[1]), dtype=np.double)
        sum_weight = np.zeros((n_imag_chan, n_imag_pol), dtype=np.double)
    do_psf = grid_params["do_psf"]
    standard_grid_jit(grid, sum_weight, do_psf, False, vis_data, uvw, freq_chan, chan_map,
                      pol_map, weight, cgk_1D, n_uv, delta_lm, support, oversampling)
    return grid, sum_weight


def standard_grid_numpy_wrap_input_checked(vis_data, grid, sum_weight, uvw, weight, freq_chan,
                                           cgk_1D, image_size, cell_size, oversampling=100,
                                           support=7, complex_grid=True, do_psf=False,
                                           chan_mode="continuum") -> Tuple[np.ndarray, np.ndarray]:
    """standard_grid.py:15-85: keyword form of standard_grid_numpy_wrap."""
    params = {
        "image_size_padded": image_size,
        "cell_size": cell_size,
        "complex_grid": complex_grid,
        "oversampling": oversampling,
        "support": support,
        "do_psf": do_psf,
        "chan_mode": chan_mode,
    }
    return standard_grid_numpy_wrap(vis_data, grid, sum_weight, uvw, weight, freq_chan, cgk_1D,
                                    params)


def standard_grid_psf_numpy_wrap(uvw, weight, freq_chan, cgk_1D, grid_params):
    """standard_grid.py:168-239: real-valued weight grid (PSF)."""
    n_chan = weight.shape[2]
    if grid_params["chan_mode"] == "cube":
        n_imag_chan, chan_map = n_chan, np.arange(n_chan)
    else:
        n_imag_chan, chan_map = 1, np.zeros(n_chan, dtype=int)
    n_imag_pol = weight.shape[3]
    n_uv = grid_params["image_size_padded"]
    grid = np.zeros((n_imag_chan, n_imag_pol, n_uv[0], n_uv[1]), dtype=np.double)
    sum_weight = np.zeros((n_imag_chan, n_imag_pol), dtype=np.double)
    standard_grid_jit(grid, sum_weight, grid_params["do_psf"],
                      grid_params["do_imaging_weight"], np.zeros((1, 1, 1, 1), dtype=bool), uvw,
                      freq_chan, chan_map, np.arange(n_imag_pol), weight, cgk_1D, n_uv,
                      grid_params["cell_size"], grid_params["support"],
                      grid_params["oversampling"])
    return grid, sum_weight


def grid2image_spheroid_ms4(vis, resid_array: np.ndarray, pixelincr: np.ndarray,
                            support: int = 7, sampling: int = 100, dopsf: bool = False,
                            column: str = "VISIBILITY", chan_mode: str = "continuum"):
    """standard_grid.py:607-708: grid one MSv4 dataset (or a list of them)
    and write the dirty image (or PSF with dopsf=True) into resid_array.

    As in the reference, WEIGHT is zeroed in place where FLAG is set or where
    UVW is NaN.  The GPU gridder masks those samples itself; the host
    mutation is kept because callers may rely on it.
    """
    listvis = datasets(vis)
    nx, ny = resid_array.shape[-2:]
    dev = _dev.require_gpu()
    cgk = _dev.ps_kernel_1d(int(sampling), int(support), dev)
    prec = _prec(np.complex128)
    gridvis = None
    sumwt = None
    gshape = None
    for elvis in listvis:
        if not is_datatree(elvis):
            raise TypeError("One of the elements of vis is not an xarray datatree")
        vis_data = elvis[column].data
        uvw = elvis.UVW.data
        weight = elvis.WEIGHT.data
        flag = elvis.FLAG.data
        weight[flag] = 0.0
        nan_uvw = np.isnan(uvw[:, :, 0]) | np.isnan(uvw[:, :, 1])
        weight[nan_uvw] = 0.0
        freq_chan = elvis.coords["frequency"].values
        n_chan, npol = weight.shape[2], weight.shape[3]
        if chan_mode == "cube":
            n_img_chan, chan_map = n_chan, np.arange(n_chan)
        else:
            n_img_chan, chan_map = 1, np.zeros(n_chan, dtype=int)
        shape = (n_img_chan, npol, nx, ny)
        if gridvis is not None and shape != gshape:
            raise RuntimeError(f"Shapes of gridded vis and image do no match {gshape} "
                               f"and {shape}")
        gridvis, sumwt = grid_to_device(vis_data, uvw, weight, flag, freq_chan, chan_map,
                                        n_img_chan, np.array([nx, ny], dtype=int), pixelincr,
                                        int(support), int(sampling), dopsf, prec, cgk,
                                        gridvis, sumwt)
        gshape = shape
        if gshape != tuple(resid_array.shape):
            raise RuntimeError(f"Shapes of gridded vis and image do no match {gshape} and "
                               f"{resid_array.shape}")
    if gridvis is None:
        return
    img = _dev.dirty_image(gridvis, sumwt, int(support), int(sampling),
                           out_dtype=torch.float64)
    resid_array[...] = img.cpu().numpy()
