"""Airy-disk aperture patterns on the GPU.

Mirror of the two functions of core/imaging/imaging_utils/make_pb_symmetric.py
that the convolution functions of the mosaic gridder use,
``airy_disk_rorder`` (:148-212) and ``casa_airy_disk_rorder`` (:215-280):
the same arguments and result ([n_antenna_type, n_freq, n_pol, n0, n1]
float64, the voltage pattern for ipower 1 and the power pattern for 2),
evaluated by ``av_gcf_aperture_patterns`` of libavgrid.so.
"""
from __future__ import annotations

import numpy as np
import torch

from . import device as _dev
from ._lib import call

C_LIGHT = 299792458.0


def aperture_patterns_device(freq_chan, pb_params, grid_params, casa: bool,
                             device=None) -> torch.Tensor:
    """[n_antenna_type, n_freq, n0, n1] float64 patterns on the device."""
    dev = _dev.require_gpu(device)
    cell = np.asarray(grid_params["cell_size"], dtype=np.float64)
    size = np.asarray(grid_params["image_size"]).astype(int)
    ctr = np.asarray(grid_params["image_center"]).astype(int)
    dishes = np.asarray(pb_params["list_dish_diameters"], dtype=np.float64)
    blocks = np.asarray(pb_params["list_blockage_diameters"], dtype=np.float64)
    if dishes.shape != blocks.shape or dishes.ndim != 1 or dishes.size == 0:
        raise ValueError("list_dish_diameters and list_blockage_diameters must be equal-length "
                         "lists")
    ipower = int(pb_params["ipower"])
    # k = (2 pi f) / c (make_pb_symmetric.py:166)
    kf = (2 * np.pi * np.asarray(freq_chan, dtype=np.float64).reshape(-1)) / C_LIGHT
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    kf_t, d_t, b_t = t(kf), t(dishes), t(blocks)
    out = torch.empty((len(dishes), len(kf), int(size[0]), int(size[1])), dtype=torch.float64,
                      device=dev)
    call("av_gcf_aperture_patterns", _dev._ptr(kf_t), len(kf), _dev._ptr(d_t), _dev._ptr(b_t),
         len(dishes), int(size[0]), int(size[1]), int(ctr[0]), int(ctr[1]), float(cell[0]),
         float(cell[1]), ipower, 1 if casa else 0, _dev._ptr(out), _dev._stream())
    return out


def _rorder(freq_chan, pol, pb_params, grid_params, casa):
    p = aperture_patterns_device(freq_chan, pb_params, grid_params, casa).cpu().numpy()
    return np.tile(p[:, :, None], (1, 1, len(pol), 1, 1))


def airy_disk_rorder(freq_chan, pol, pb_params, grid_params):
    """make_pb_symmetric.py:148-212: obscured airy pattern per antenna type."""
    return _rorder(freq_chan, pol, pb_params, grid_params, False)


def casa_airy_disk_rorder(freq_chan, pol, pb_params, grid_params):
    """make_pb_symmetric.py:215-280: CASA's obscured airy pattern
    (PBMath1DAiry) per antenna type."""
    return _rorder(freq_chan, pol, pb_params, grid_params, True)
