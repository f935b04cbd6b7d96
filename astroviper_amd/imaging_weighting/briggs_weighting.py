"""Briggs weighting factors (mirror of
core/imaging/imaging_weighting/briggs_weighting.py:1-31), with the density
grid's sum of squares formed on the GPU."""
from __future__ import annotations

import numpy as np
import torch

from .. import device as _dev


def calculate_briggs_params(grid_of_imaging_weights, sum_weight, imaging_weights_parms):
    """Factors (a, b) of the Briggs denominator a * G + b per (chan, pol):
    "briggs": a = (5 * 10^-robust)^2 / (sum G^2 / sum_weight), b = 1;
    "briggs_abs": a = robust^2, b = 2 * briggs_abs_noise^2; otherwise a = 1,
    b = 0 (natural)."""
    kind = imaging_weights_parms["weighting"]
    if kind == "briggs":
        robust = imaging_weights_parms["robust"]
        bf = np.ones((2,) + sum_weight.shape)
        dev = _dev.require_gpu()
        g = torch.from_numpy(np.ascontiguousarray(grid_of_imaging_weights, dtype=np.float64))
        sq = torch.sum(g.to(dev) ** 2, dim=(2, 3)).cpu().numpy()
        with np.errstate(invalid="ignore", divide="ignore"):
            bf[0, :, :] = (np.square(5.0 * 10.0 ** (-robust)) / (sq / sum_weight))[None, None]
    elif kind == "briggs_abs":
        robust = imaging_weights_parms["robust"]
        bf = np.ones((2,) + sum_weight.shape)
        bf[0, :, :] = bf[0, :, :] * np.square(robust)
        bf[1, :, :] = bf[1, :, :] * 2.0 * np.square(imaging_weights_parms["briggs_abs_noise"])
    else:
        bf = np.zeros((2, 1, 1) + sum_weight.shape)
        bf[0, :, :] = np.ones((1, 1, 1) + sum_weight.shape)
    return bf
