"""make_psf on the GPU (mirror of
core/imaging/imaging_utils/make_point_spread_function.py:1-79)."""
from __future__ import annotations

import numpy as np

from .msv4 import DataArray, ImageDataset
from .standard_grid import grid2image_spheroid_ms4


def _sky_image(phase_center, image_size, cell_size, freq, pols, time):
    """Image dataset with the (time, frequency, polarization, l, m)
    coordinates of an xradio sky image (SIN projection offsets in radians)."""
    nx, ny = int(image_size[0]), int(image_size[1])
    l = (np.arange(nx) - nx // 2) * (-float(cell_size[0]))
    m = (np.arange(ny) - ny // 2) * float(cell_size[1])
    coords = {"time": np.atleast_1d(time), "frequency": np.asarray(freq),
              "polarization": np.asarray(pols), "l": l, "m": m}
    return ImageDataset({}, coords, attrs={"phase_center": phase_center,
                                           "projection": "SIN",
                                           "direction_reference": "fk5"})


def make_psf(vis, im_params, grid_params):
    """Grid the weights of `vis` and return an image dataset holding
    POINT_SPREAD_FUNCTION (time, frequency, polarization, l, m).

    im_params: cell_size (rad, 2), image_size (nx, ny), phase_center,
    chan_mode; grid_params: sampling, support (make_point_spread_function.py:1-22).
    """
    freq_chan = vis.coords["frequency"].values
    nfreq = len(freq_chan)
    image_size = im_params["image_size"]
    cell_size = im_params["cell_size"]
    phase_center = im_params["phase_center"]
    chan_mode = im_params["chan_mode"]
    sampling = grid_params["sampling"]
    support = grid_params["support"]
    time_coords = vis.coords["time"].values[0]
    pol = vis.coords["polarization"].values
    npol = len(pol)
    psf_data = np.zeros([nfreq, npol, image_size[1], image_size[0]], dtype=float)
    grid2image_spheroid_ms4(vis=vis, resid_array=psf_data,
                            pixelincr=np.array([-cell_size[0], cell_size[1]]),
                            support=support, sampling=sampling, dopsf=True,
                            chan_mode=chan_mode)
    psf_xds = _sky_image(phase_center, image_size, cell_size, freq_chan, pol, time_coords)
    psf_xds["POINT_SPREAD_FUNCTION"] = DataArray(
        np.expand_dims(psf_data, axis=0), dims=("time", "frequency", "polarization", "l", "m"),
        coords=psf_xds.coords, name="POINT_SPREAD_FUNCTION")
    return psf_xds
