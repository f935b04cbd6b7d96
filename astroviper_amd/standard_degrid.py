"""Standard degridder on the GPU.

Mirror of core/imaging/imaging_utils/standard_degrid.py: degrid_spheroid_ms4
and the dgrid_numba kernel signature, computed by libavgrid.so.  The
reference's alternative kernels (whichFunc 1 = dgrid_optimized, 2 = dgrid)
compute the same values; every whichFunc runs the one GPU degridder.
"""
from __future__ import annotations

import numpy as np
import torch

from . import device as _dev
from .msv4 import make_data_array
from .standard_grid import _MAX_PLAN_SAMPLES, _cgk_device, _prec

C_LIGHT = 299792458.0  # astropy.constants.c


def _degrid_to_device(uvw, flag, freq, grid, scale, offset, support, sampling, cgk_dev,
                      chanmap, polmap, values_dev, incremental, tile=(16, 16)):
    """Degrid host uvw [nt, nb, 3] / flag [nt, nb, nf, npol] against the grid
    into values_dev [rows, nf, npol] on the GPU."""
    dev = _dev.require_gpu()
    flag = np.asarray(flag)
    nt, nb, nf, npol = flag.shape
    rows = nt * nb
    g = np.asarray(grid)
    prec = _prec(g.dtype)
    cdt = torch.complex64 if prec == _dev.AV_F32 else torch.complex128
    g_dev = torch.from_numpy(np.ascontiguousarray(g)).to(dev).to(cdt)
    nchan_g, npol_g, n_u, n_v = g.shape
    uvw_h = np.ascontiguousarray(np.asarray(uvw, dtype=np.float64).reshape(rows, 3))
    f_h = np.ascontiguousarray(flag.reshape(rows, nf, npol)).astype(np.uint8)
    freq_t = torch.from_numpy(np.ascontiguousarray(freq, dtype=np.float64)).to(dev)
    cmap = np.asarray(chanmap).astype(np.int32)
    # channels mapped outside the grid are skipped (standard_degrid.py:423)
    cmap = np.where((cmap >= 0) & (cmap < nchan_g), cmap, -1).astype(np.int32)
    cmap_t = torch.from_numpy(cmap).to(dev)
    # polarisations mapped outside the grid are skipped by av_degrid itself
    # (standard_degrid.py:419-420)
    pmap_t = torch.from_numpy(np.asarray(polmap).astype(np.int32)).to(dev)
    # blocks of whole time steps, so that each plan's rows are time-major
    step = max(1, _MAX_PLAN_SAMPLES // max(1, nf) // max(1, nb)) * max(1, nb)
    for r0 in range(0, rows, step):
        r1 = min(rows, r0 + step)
        # NaN cells: the plan computes no gridding geometry, only the
        # degridder's (with the transform below)
        plan = _dev.GridPlan(r1 - r0, nf, nchan_g, (n_u, n_v), (np.nan, np.nan), support,
                             sampling, tile=tile, device=dev, n_baseline=max(1, nb))
        plan.cgk = cgk_dev
        _dev.call("av_plan_set_degrid_transform", plan._h, float(scale[0]), float(scale[1]),
                  float(offset[0]), float(offset[1]))
        plan.set_uvw(torch.from_numpy(uvw_h[r0:r1]).to(dev), freq_t, cmap_t)
        fl = f_h[r0:r1]
        out = values_dev[r0:r1].reshape((r1 - r0) * nf, npol)
        plan.degrid(g_dev, torch.from_numpy(fl).to(dev), out, n_pol=npol, pol_map=pmap_t,
                    incremental=incremental)
        plan.close()
    return values_dev


def dgrid_numba(uvw, dphase, values, flag, scale, offset, grid, freq, c, support, sampling,
                convFunc, chanmap, polmap):
    """standard_degrid.py:388-485 on the GPU: values += degridded samples,
    in place.  Only zero dphase (the value degrid_spheroid_ms4 passes) is
    supported."""
    if np.any(np.asarray(dphase) != 0):
        raise NotImplementedError("non-zero dphase (phase rotation) is not implemented")
    if float(c) != C_LIGHT:
        raise NotImplementedError("c must be the speed of light, 299792458.0")
    dev = _dev.require_gpu()
    prec = _prec(np.asarray(values).dtype)
    cdt = torch.complex64 if prec == _dev.AV_F32 else torch.complex128
    cgk = _cgk_device(convFunc, int(sampling), int(support), dev)
    nt, nb, nf, npol = values.shape
    vals = torch.from_numpy(np.ascontiguousarray(values)).to(dev).to(cdt)
    vals = vals.reshape(nt * nb, nf, npol).contiguous()
    g = np.asarray(grid)
    if _prec(g.dtype) != prec:
        g = g.astype(np.complex64 if prec == _dev.AV_F32 else np.complex128)
    _degrid_to_device(uvw, flag, freq, g, scale, offset, int(support), int(sampling), cgk,
                      chanmap, polmap, vals, incremental=True)
    values[...] = vals.reshape(nt, nb, nf, npol).cpu().numpy()


def degrid_spheroid_ms4(vis, grid: np.ndarray, pixelincr: np.ndarray, support: int = 7,
                        sampling: int = 100, incremental: bool = False, whichFunc=0):
    """standard_degrid.py:488-608: predict VISIBILITY_MODEL from a model grid.

    As in the reference, rows with NaN UVW are flagged and their UVW set to
    1e9 in place, and vis["VISIBILITY_MODEL"] is set (complex128).
    """
    uvw = vis.UVW.data
    dims = vis.dims
    nt, nb, nf, npl = dims["time"], dims["baseline_id"], dims["frequency"], dims["polarization"]
    if not incremental or "VISIBILITY_MODEL" not in vis:
        modvis = np.zeros((nt, nb, nf, npl), dtype=np.complex128)
    else:
        modvis = vis.VISIBILITY_MODEL.data
    flag = vis.FLAG.data
    nan_rows = np.isnan(uvw).any(axis=2)
    if nan_rows.any():
        it, ib = np.nonzero(nan_rows)
        flag[it, ib, :, :] = True
        uvw[it, ib, :] = 1e9
    nx, ny = grid.shape[-2:]
    scale = np.array([-pixelincr[0] * nx, -pixelincr[1] * ny, 0.0]).astype(np.float64)
    offset = np.array([nx / 2, ny / 2, 0.0]).astype(np.float64)
    freq = vis.frequency.data
    nchan = grid.shape[0]
    npol = grid.shape[1]
    chanmap = (np.arange(nf) / nf * nchan).astype(int)
    polmap = np.round(np.arange(npl) % npol).astype(int)
    dev = _dev.require_gpu()
    # modvis is complex128, so the sum is formed in FP64 as in the reference
    g = np.asarray(grid).astype(np.complex128, copy=False)
    cgk = _dev.ps_kernel_1d(int(sampling), int(support), dev)
    vals = torch.from_numpy(np.ascontiguousarray(modvis)).to(dev).to(torch.complex128)
    vals = vals.reshape(nt * nb, nf, npl).contiguous()
    _degrid_to_device(uvw, flag, freq, g, scale, offset, int(support), int(sampling), cgk,
                      chanmap, polmap, vals, incremental=True)
    modvis[...] = vals.reshape(nt, nb, nf, npl).cpu().numpy()
    datdims = vis.VISIBILITY.dims
    datcoords = vis.VISIBILITY.coords
    vis["VISIBILITY_MODEL"] = make_data_array(modvis, coords=datcoords, dims=datdims)
