"""Direct callers of the gridder in core/imaging/imager.py:54-105."""
from __future__ import annotations

from typing import Tuple

import numpy as np

from .fft import fft_lm_to_uv
from .msv4 import make_data_array
from .standard_degrid import degrid_spheroid_ms4
from .standard_grid import grid2image_spheroid_ms4


def grid_visibilities(ms4, image_size: Tuple[int, int], cell_size: Tuple[float, float],
                      support: int = 7, oversampling: int = 100, chan_mode: str = "cube",
                      column: str = "VISIBILITY") -> np.ndarray:
    """Grid visibilities to the image plane; returns (n_chan, n_pol, ny, nx)
    (imager.py:54-79)."""
    n_chan = len(ms4.coords["frequency"])
    n_pol = len(ms4.coords["polarization"])
    ny, nx = image_size
    dirty_image = np.zeros([n_chan, n_pol, ny, nx], dtype=float)
    grid2image_spheroid_ms4(vis=ms4, resid_array=dirty_image, pixelincr=np.array(cell_size),
                            support=support, sampling=oversampling, dopsf=False,
                            column=column, chan_mode=chan_mode)
    return dirty_image


def compute_residual_visibilities(ms4) -> None:
    """RESIDUAL = VISIBILITY - VISIBILITY_MODEL in place (imager.py:82-85)."""
    v = ms4["VISIBILITY"]
    ms4["RESIDUAL"] = make_data_array(v.data - ms4["VISIBILITY_MODEL"].data, coords=v.coords,
                                      dims=v.dims)


def predict_model_visibilities(ms4, model_image: np.ndarray, cell_size: Tuple[float, float],
                               support: int = 7, oversampling: int = 100,
                               incremental: bool = False) -> None:
    """FFT the model image to the uv plane (no taper, as in the reference)
    and degrid into VISIBILITY_MODEL (imager.py:88-105)."""
    model_grid = fft_lm_to_uv(model_image, axes=[2, 3])
    degrid_spheroid_ms4(vis=ms4, grid=model_grid, pixelincr=np.array(cell_size),
                        support=support, sampling=oversampling, incremental=incremental)
