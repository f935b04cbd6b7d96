"""xradio-image level gridding (mirror of
core/imaging/imaging_utils/standard_image_grid_preparation.py:312-580).

Image datasets carry VISIBILITY / UV_SAMPLING (time, frequency, polarization,
u, v) complex planes and their *_NORMALIZATION sums of weights; l/m images
carry SKY / RESIDUAL / MODEL.  FFTs run on the GPU.
"""
from __future__ import annotations

from typing import List, Union

import numpy as np

from .fft import fft_lm_to_uv
from .gcf_prolate_spheroidal import create_prolate_spheroidal_kernel
from .ifft import ifft_uv_to_lm
from .msv4 import datasets, is_datatree
from .standard_grid import standard_grid_numpy_wrap_input_checked
from .gcf_prolate_spheroidal import create_prolate_spheroidal_kernel_1D


def _vars(data_vars, default):
    if not data_vars:
        return default
    if isinstance(data_vars, str):
        return [data_vars]
    return data_vars


def fft_to_uv(image, data_vars: Union[str, List[str]] = "") -> None:
    """standard_image_grid_preparation.py:312-341."""
    for sky_var in _vars(data_vars, ["SKY", "RESIDUAL", "MODEL"]):
        if sky_var in image.data_vars:
            dat_name = "VISIBILITY_" + sky_var
            if dat_name not in image.data_vars:
                raise Exception(f"{dat_name} is not in {image}")
            image[dat_name].data[...] = fft_lm_to_uv(image[sky_var].data, axes=[3, 4])


def ifft_to_lm(image, data_vars: Union[str, List[str]] = "") -> None:
    """standard_image_grid_preparation.py:344-377."""
    for vis_var in _vars(data_vars, ["VISIBILITY", "VISIBILITY_RESIDUAL", "VISIBILITY_MODEL"]):
        if vis_var in image.data_vars:
            dat_name = "SKY" if vis_var == "VISIBILITY" else vis_var.split("VISIBILITY_", 1)[1]
            image[dat_name].data[...] = ifft_uv_to_lm(image[vis_var].data, axes=[3, 4])


def correct_fft_to_uv(image, data_vars: Union[str, List[str]] = "") -> None:
    """standard_image_grid_preparation.py:380-409: divide by u_size * v_size."""
    u_size = image.sizes["u"]
    v_size = image.sizes["v"]
    for vis_var in _vars(data_vars, ["VISIBILITY", "VISIBILITY_RESIDUAL", "VISIBILITY_MODEL"]):
        if vis_var in image.data_vars:
            image[vis_var].data /= u_size * v_size


def correct_ifft_to_lm(image, data_vars: Union[str, List[str]] = "", doSpheroidCorr: bool = False,
                       convsampling: int = 100, convsupport: int = 7) -> None:
    """standard_image_grid_preparation.py:412-492: multiply by nx*ny, or with
    doSpheroidCorr by nx*ny / (taper * sum of weights) per plane."""
    nx = image.sizes["l"]
    ny = image.sizes["m"]
    imsize = np.array([nx, ny])
    corr = None
    if doSpheroidCorr:
        _, corr = create_prolate_spheroidal_kernel(convsampling, convsupport, imsize)
    for sky_var in _vars(data_vars, ["SKY", "RESIDUAL"]):
        if sky_var not in image.data_vars:
            continue
        if not doSpheroidCorr:
            image[sky_var].data[...] *= nx * ny
            continue
        sumwt = ("VISIBILITY_NORMALIZATION" if sky_var == "SKY"
                 else "VISIBILITY_" + sky_var + "_NORMALIZATION")
        if sumwt not in image.data_vars:
            raise Exception(f"Do not have {sumwt} to do image correction")
        dat = image[sky_var].data
        sw = image[sumwt].data
        for t in range(dat.shape[0]):
            for f in range(dat.shape[1]):
                for p in range(dat.shape[2]):
                    dat[t, f, p] *= (nx * ny) / (corr * sw[t, f, p])


def grid2xradio_spheroid_ms4(vis, image, support: int = 7, sampling: int = 100,
                             data_var: str = "VISIBILITY", chan_mode: str = "continuum"):
    """standard_image_grid_preparation.py:495-580: accumulate the gridded
    data (or weights for UV_SAMPLING) of each time plane into image[data_var]
    and image[data_var + "_NORMALIZATION"] in place."""
    listvis = datasets(vis)
    nx = image.sizes["u"]
    ny = image.sizes["v"]
    dopsf = data_var == "UV_SAMPLING"
    cellsize_l = np.abs(image.l[1].data - image.l[0].data)
    cellsize_m = np.abs(image.m[1].data - image.m[0].data)
    pixelincr = np.array([cellsize_l, cellsize_m])
    if data_var not in image.data_vars:
        raise Exception(f"{data_var} is not in {image}")
    sumwt_var = data_var + "_NORMALIZATION"
    dat = image[data_var].data
    sw_all = image[sumwt_var].data
    cgk_1D = create_prolate_spheroidal_kernel_1D(sampling, support)
    for ti in range(dat.shape[0]):
        gridvis = dat[ti]
        sumwt = sw_all[ti]
        for elvis in listvis:
            if not is_datatree(elvis):
                raise TypeError("One of the elements of vis is not an xarray datatree")
            vis_data = elvis[data_var].data if not dopsf else None
            uvw = elvis.UVW.data
            weight = elvis.WEIGHT.data
            flag = elvis.FLAG.data
            weight[flag] = 0.0
            freq_chan = elvis.coords["frequency"].values
            gridvis, sumwt = standard_grid_numpy_wrap_input_checked(
                vis_data, gridvis, sumwt, uvw, weight, freq_chan, cgk_1D,
                image_size=np.array([nx, ny], dtype=int), cell_size=pixelincr,
                oversampling=sampling, support=support, complex_grid=True, do_psf=dopsf,
                chan_mode=chan_mode)
