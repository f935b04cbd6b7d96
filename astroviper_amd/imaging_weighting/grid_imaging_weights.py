"""Imaging-weight gridding on the GPU: mirror of
core/imaging/imaging_weighting/grid_imaging_weights.py with the same entry
points, arguments, in-place semantics and results, computed by libavgrid.so
(av_grid_imaging_weights / av_degrid_imaging_weights).

One deliberate difference: the reference does not bound the conjugate pixel
of a sample, which lies one past the end of the grid for a sample on row or
column 0 of an even-sized grid; that one update is skipped here.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _lib
from .. import device as _dev


def _t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)


def grid_imaging_weights(grid: np.ndarray, sum_weight: np.ndarray, uvw: np.ndarray,
                         data_weight: np.ndarray, freq_chan: np.ndarray, grid_parms: dict):
    """grid_imaging_weights.py:7-100: add the data weights (polarisation 0)
    of every sample at its nearest pixel and the conjugate pixel of
    grid (n_chan, n_pol, n_u, n_v), and 2 * weight to sum_weight
    (n_chan, n_pol), both in place (plane and column 0)."""
    assert data_weight.shape[3] < 3, "Polarization should be PP or PP, QQ."
    dev = _dev.require_gpu()
    n_uv = grid_parms["image_size"]
    cell = grid_parms["cell_size"]
    nt, nb, nchan, npol = data_weight.shape
    n_u, n_v = int(n_uv[0]), int(n_uv[1])
    g = _t(grid[:, 0], dev)
    sw = _t(sum_weight[:, 0], dev)
    u = _t(np.asarray(uvw).reshape(nt * nb, 3), dev)
    w = _t(data_weight, dev)
    f = _t(freq_chan, dev)
    _lib.call("av_grid_imaging_weights", _dev._ptr(u), nt * nb, nb, _dev._ptr(f), nchan,
              _dev._ptr(w), npol, n_u, n_v, float(cell[0]), float(cell[1]), _dev._ptr(g),
              _dev._ptr(sw), _dev._stream())
    grid[:, 0] = g.cpu().numpy()
    sum_weight[:, 0] = sw.cpu().numpy()


def degrid_imaging_weights(grid_imaging_weight, uvw, data_weight, briggs_factors, freq_chan,
                           grid_parms):
    """grid_imaging_weights.py:209-271: per-visibility imaging weights
    (shape of data_weight, single polarisation): the data weight, divided by
    briggs_factors[0] * G + briggs_factors[1] with G the weight-density grid
    at the sample's pixel; 0 outside the grid."""
    dev = _dev.require_gpu()
    n_uv = grid_parms["image_size"]
    cell = grid_parms["cell_size"]
    nt, nb, nchan, npol = data_weight.shape
    n_u, n_v = int(n_uv[0]), int(n_uv[1])
    g = _t(np.asarray(grid_imaging_weight)[:, 0], dev)
    bf = np.asarray(briggs_factors, dtype=np.float64).reshape(2, -1)
    b = _t(np.stack([np.broadcast_to(bf[0], (nchan,)), np.broadcast_to(bf[1], (nchan,))]), dev)
    u = _t(np.asarray(uvw).reshape(nt * nb, 3), dev)
    w = _t(np.asarray(data_weight)[..., 0], dev)
    f = _t(freq_chan, dev)
    out = torch.empty((nt * nb, nchan), dtype=torch.float64, device=dev)
    flag = torch.zeros(1, dtype=torch.int32, device=dev)
    _lib.call("av_degrid_imaging_weights", _dev._ptr(g), _dev._ptr(b), _dev._ptr(u), nt * nb, nb,
              _dev._ptr(f), nchan, _dev._ptr(w), n_u, n_v, float(cell[0]), float(cell[1]),
              _dev._ptr(out), _dev._ptr(flag), _dev._stream())
    if int(flag.item()):
        raise ValueError("NaN in briggs_factors")
    res = np.zeros(data_weight.shape, dtype=np.double)
    res[..., 0] = out.cpu().numpy().reshape(nt, nb, nchan)
    return res
