"""Mosaic / A-projection gridder on the GPU.

Mirror of core/imaging/imaging_utils/mosaic_grid.py: ``mosaic_grid_jit`` with
the reference's arguments and in-place accumulation, run by
``av_mosaic_grid`` of libavgrid.so over a plan of the samples.  The
reference's two wrappers call their kernels with arguments the kernels do
not take (``mosaic_grid_numpy_wrap`` calls ``aperture_grid_jit`` and
``mosaic_psf_grid_numpy_wrap`` calls ``mosaic_grid_jit`` with ``field`` and
``field_id``, which it lacks, and both use ``np.int``), so they cannot run;
here they prepare the same arguments and call ``mosaic_grid_jit``.
"""
from __future__ import annotations

import numpy as np
import torch

from . import device as _dev
from ._lib import AV_F32, AV_F64, AV_GRID_ACCUMULATE, AV_GRID_PSF, call


def _tile_for(support: int) -> int:
    for t in (16, 32, 64, 128):
        if t >= support:
            return t
    raise NotImplementedError(f"convolution support {support} exceeds the widest tile")


def _plan_shape(max_support: int):
    """The binning plan's support and tile for a mosaic call.  The plan
    records each sample in the tiles its footprint (at most max_support taps
    per axis) touches, and its bounds test at that support keeps every sample
    the reference keeps; the gridder applies the reference's own test
    (centre +- max_support inside the grid, mosaic_grid.py:379-383)."""
    return max(2, max_support), _tile_for(2 * max_support + 1)


def mosaic_grid_jit(grid, sum_weight, do_psf, vis_data, uvw, freq_chan, chan_map, pol_map,
                    cf_baseline_map, cf_chan_map, cf_pol_map, imaging_weight, conv_kernel, n_uv,
                    delta_lm, weight_support, oversampling, phase_gradient):
    """mosaic_grid.py:306-486 on the GPU: accumulate into the complex `grid`
    [n_chan, n_pol, n_u, n_v] and `sum_weight` [n_chan, n_pol] in place.
    complex64 grids are gridded in FP32, complex128 in FP64."""
    if not np.iscomplexobj(grid):
        raise TypeError("the mosaic gridder needs a complex grid")
    iw = np.asarray(imaging_weight)
    nt, nb, nchan, npol = iw.shape
    if not np.array_equal(np.asarray(pol_map), np.arange(npol)) or grid.shape[1] != npol:
        raise NotImplementedError("pol_map must be arange(imaging_weight.shape[3])")
    fp32 = grid.dtype == np.complex64
    rdt = torch.float32 if fp32 else torch.float64
    cdt = torch.complex64 if fp32 else torch.complex128
    dev = _dev.require_gpu()
    ws = np.ascontiguousarray(weight_support, dtype=np.int32)
    max_support = int(np.max(ws))
    support, tile = _plan_shape(max_support)
    ov = (int(oversampling[0]), int(oversampling[1]))
    ck = np.asarray(conv_kernel)
    ncb, ncc, ncp, cu_n, cv_n = ck.shape
    # conv_kernel * phase_gradient[0] (mosaic_grid.py:359), on the device
    cfpg = (torch.from_numpy(np.ascontiguousarray(ck, dtype=np.complex128)).to(dev)
            * torch.from_numpy(np.ascontiguousarray(np.asarray(phase_gradient)[0],
                                                    dtype=np.complex128)).to(dev)).to(cdt)
    maps = (np.asarray(cf_baseline_map), np.asarray(cf_chan_map), np.asarray(cf_pol_map))
    for m, k, lim, what in ((maps[0], nb, ncb, "cf_baseline_map"), (maps[1], nchan, ncc,
                            "cf_chan_map"), (maps[2], npol, ncp, "cf_pol_map")):
        if m.size < k or np.any(m[:k] < 0) or np.any(m[:k] >= lim):
            raise ValueError(f"{what} needs {k} entries in [0, {lim})")
    if ws.shape != (ncb, ncc, ncp, 2) or np.any(ws < 1):
        raise ValueError("weight_support must be [n_cf_baseline, n_cf_chan, n_cf_pol, 2] >= 1")
    if min(ov) < 1 or max(ov) > 254:
        raise ValueError("oversampling must lie in [1, 254]")
    rows = nt * nb
    plan = _dev.GridPlan(rows, nchan, grid.shape[0], n_uv, delta_lm, support,
                         max(ov[0], 1), tile=(tile, tile), device=dev, n_baseline=nb)
    try:
        plan.set_uvw(torch.from_numpy(np.ascontiguousarray(uvw, dtype=np.float64)
                                      .reshape(rows, 3)).to(dev),
                     torch.from_numpy(np.ascontiguousarray(freq_chan, dtype=np.float64)).to(dev),
                     torch.from_numpy(np.ascontiguousarray(chan_map, dtype=np.int32)).to(dev))
        i32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(dev)
        g_dev = torch.from_numpy(np.ascontiguousarray(grid)).to(dev)
        s_dev = torch.from_numpy(np.ascontiguousarray(sum_weight, dtype=np.float64)).to(dev)
        w_dev = torch.from_numpy(np.ascontiguousarray(iw)).to(dev, rdt)
        v_dev = None
        if not do_psf:
            v_dev = torch.from_numpy(np.ascontiguousarray(
                np.asarray(vis_data)[..., :npol])).to(dev).to(cdt)
        keep = (i32(maps[0][:nb]), i32(maps[1][:nchan]), i32(maps[2][:npol]), i32(ws))
        mode = AV_GRID_ACCUMULATE | (AV_GRID_PSF if do_psf else 0)
        p = _dev._ptr
        call("av_mosaic_grid", plan._h, p(v_dev), p(w_dev), int(npol), int(nb), p(keep[0]),
             p(keep[1]), p(keep[2]), p(cfpg), int(ncb), int(ncc), int(ncp), int(cu_n), int(cv_n),
             p(keep[3]), ov[0], ov[1], p(g_dev), p(s_dev), AV_F32 if fp32 else AV_F64, mode,
             _dev._stream())
        grid[...] = g_dev.cpu().numpy()
        sum_weight[...] = s_dev.cpu().numpy()
    finally:
        plan.close()
    return grid, sum_weight


def _maps(n_chan, npol, chan_mode):
    if chan_mode == "cube":
        return n_chan, np.arange(n_chan)
    return 1, np.zeros(n_chan, dtype=int)


def mosaic_grid_numpy_wrap(vis_data, uvw, imaging_weight, field, cf_baseline_map, cf_chan_map,
                           cf_pol_map, conv_kernel, weight_support, phase_gradient, freq_chan,
                           grid_params):
    """mosaic_grid.py:156-227 (see the module notes): allocate the grid and
    sum of weights and grid the visibilities."""
    n_chan, npol = imaging_weight.shape[2], imaging_weight.shape[3]
    n_img_chan, chan_map = _maps(n_chan, npol, grid_params["chan_mode"])
    n_uv = grid_params["image_size_padded"]
    grid = np.zeros((n_img_chan, npol, n_uv[0], n_uv[1]), dtype=np.complex128)
    sum_weight = np.zeros((n_img_chan, npol), dtype=np.double)
    mosaic_grid_jit(grid, sum_weight, grid_params["do_psf"], vis_data, uvw, freq_chan, chan_map,
                    np.arange(npol), cf_baseline_map, cf_chan_map, cf_pol_map, imaging_weight,
                    conv_kernel, np.asarray(n_uv), grid_params["cell_size"], weight_support,
                    grid_params["oversampling"], phase_gradient)
    return grid, sum_weight


def mosaic_psf_grid_numpy_wrap(uvw, imaging_weight, field, cf_baseline_map, cf_chan_map,
                               cf_pol_map, conv_kernel, weight_support, phase_gradient, freq_chan,
                               grid_params):
    """mosaic_grid.py:230-303 (see the module notes): the weight grid."""
    n_chan, npol = imaging_weight.shape[2], imaging_weight.shape[3]
    n_img_chan, chan_map = _maps(n_chan, npol, grid_params["chan_mode"])
    n_uv = grid_params["image_size_padded"]
    grid = np.zeros((n_img_chan, npol, n_uv[0], n_uv[1]), dtype=np.complex128)
    sum_weight = np.zeros((n_img_chan, npol), dtype=np.double)
    mosaic_grid_jit(grid, sum_weight, True, np.zeros((1, 1, 1, 1), dtype=bool), uvw, freq_chan,
                    chan_map, np.arange(npol), cf_baseline_map, cf_chan_map, cf_pol_map,
                    imaging_weight, conv_kernel, np.asarray(n_uv), grid_params["cell_size"],
                    weight_support, grid_params["oversampling"], phase_gradient)
    return grid, sum_weight
