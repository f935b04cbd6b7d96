"""Prolate-spheroidal gridding kernel and correction taper, computed on the GPU.

Mirror of core/imaging/imaging_utils/gcf_prolate_spheroidal.py: same names,
arguments and NumPy return values; the arithmetic runs in libavgrid.so
(Schwab's rational approximation, M=6, alpha=1).
"""
from __future__ import annotations

import numpy as np
import torch

from . import device as _dev


def prolate_spheroidal_function(u):
    """gcf_prolate_spheroidal.py:102-158 -> (grdsf, (1 - u^2) * grdsf)."""
    u_np = np.asarray(u, dtype=np.float64)
    dev = _dev.require_gpu()
    t = torch.from_numpy(np.ascontiguousarray(u_np).reshape(-1).copy()).to(dev)
    g, k = _dev.ps_function(t)
    return (g.cpu().numpy().reshape(u_np.shape), k.cpu().numpy().reshape(u_np.shape))


def create_prolate_spheroidal_kernel_1D(oversampling, support):
    """gcf_prolate_spheroidal.py:87-99: the half kernel sampled at
    i / (support//2 * oversampling), zero-padded by one oversampled pixel."""
    return _dev.ps_kernel_1d(int(oversampling), int(support)).cpu().numpy()


def coordinates(npixel: int):
    """1D array which spans [-.5,.5[ with 0 at position npixel/2
    (gcf_prolate_spheroidal.py:161-163)."""
    return (np.arange(npixel) - npixel // 2) / npixel


def coordinates2(npixel: int):
    """Two-dimensional coordinates spanning [-.5,.5[ (gcf_prolate_spheroidal.py:166-171)."""
    return (np.mgrid[0:npixel, 0:npixel] - npixel // 2) / npixel


def create_prolate_spheroidal_kernel(oversampling, support, n_uv):
    """gcf_prolate_spheroidal.py:8-84 -> (kernel, kernel_image).

    kernel is [nrow, nrow, support, support] (nrow = oversampling + 1 for even
    oversampling) with kernel[x, y] = outer(k1[x], k1[y]) for x, y <
    oversampling; kernel_image is the gridding-correction taper [n_uv[0], n_uv[1]].
    """
    dev = _dev.require_gpu()
    support_center = support // 2
    oversampling_center = oversampling // 2
    support_values = np.arange(support) - support_center
    nrow = oversampling + 1 if oversampling % 2 == 0 else oversampling
    ov_values = ((np.arange(nrow) - oversampling_center) / oversampling)[:, None]
    points = (np.broadcast_to(support_values, (nrow, support)) + ov_values) / support_center
    _, k1 = _dev.ps_function(torch.from_numpy(np.ascontiguousarray(points)).to(dev))
    kernel = _dev.ps_outer(k1.reshape(nrow, support), min(oversampling, nrow))
    n_u, n_v = int(n_uv[0]), int(n_uv[1])
    taper = torch.empty((n_u, n_v), dtype=torch.float64, device=dev)
    _dev.call("av_ps_taper", n_u, n_v, _dev._ptr(taper), _dev._stream())
    return kernel.cpu().numpy(), taper.cpu().numpy()
