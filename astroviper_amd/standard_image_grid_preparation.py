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


def make_empty_padded_uv_image(image, image_size, uv_data_array: str = "VISIBILITY_RESIDUAL",
                               insert_res_or_mod: bool = True):
    """standard_image_grid_preparation.py:123-233: an image dataset of
    `image_size` carrying (time, frequency, polarization, u, v) complex64
    zeros in `uv_data_array` and float64 zeros in its _NORMALIZATION, on the
    l/m coordinates of `image` padded to `image_size`.  With
    insert_res_or_mod, the MODEL or RESIDUAL image matching
    VISIBILITY_MODEL / VISIBILITY_RESIDUAL is copied into the centre of a
    float32 image of the padded size.  u and v are in wavelengths,
    (i - n//2) / (n * cell), so index n//2 is the origin."""
    from .msv4 import DataArray, ImageDataset
    image_size = np.array(image_size)
    cur = np.array([image.sizes["l"], image.sizes["m"]])
    if np.any(cur > image_size):
        raise Exception(f"current image size {cur} is larger than requested {image_size}")
    cl = float(image.l[1].data - image.l[0].data)
    cm = float(image.m[1].data - image.m[0].data)
    nx, ny = int(image_size[0]), int(image_size[1])
    coords = {
        "time": image.coords["time"].values,
        "frequency": image.coords["frequency"].values,
        "polarization": image.coords["polarization"].values,
        "l": (np.arange(nx) - nx // 2) * cl,
        "m": (np.arange(ny) - ny // 2) * cm,
        "u": (np.arange(nx) - nx // 2) / (nx * abs(cl)),
        "v": (np.arange(ny) - ny // 2) / (ny * abs(cm)),
    }
    lead = (len(coords["time"]), len(coords["frequency"]), len(coords["polarization"]))
    out = ImageDataset({}, coords, attrs=dict(getattr(image, "attrs", {})))
    uv_dims = ("time", "frequency", "polarization", "u", "v")
    out[uv_data_array] = DataArray(np.zeros(lead + (nx, ny), dtype=np.complex64), dims=uv_dims,
                                   coords={k: coords[k] for k in uv_dims})
    out[uv_data_array + "_NORMALIZATION"] = DataArray(
        np.zeros(lead, dtype=np.float64), dims=uv_dims[:3],
        coords={k: coords[k] for k in uv_dims[:3]})
    if insert_res_or_mod:
        name = {"VISIBILITY_MODEL": "MODEL", "VISIBILITY_RESIDUAL": "RESIDUAL"}.get(uv_data_array)
        if name and name in image.data_vars:
            dat = np.asarray(image[name].data)
            blc = (image_size - np.array(dat.shape[-2:])) // 2
            trc = blc + np.array(dat.shape[-2:])
            sky_dims = ("time", "frequency", "polarization", "l", "m")
            padded = np.zeros(lead + (nx, ny), dtype=np.float32)
            padded[:, :, :, blc[0]:trc[0], blc[1]:trc[1]] = dat
            out[name] = DataArray(padded, dims=sky_dims, coords={k: coords[k] for k in sky_dims})
    return out


def fft_to_uv(image, data_vars: Union[str, List[str]] = "") -> None:
    """standard_image_grid_preparation.py:312-341."""
    for sky_var in _vars(data_vars, ["SKY", "RESIDUAL", "MODEL"]):
        if sky_var in image.data_vars:
            dat_name = "VISIBILITY_" + sky_var
            if dat_name not in image.data_vars:
                raise Exception(f"{dat_name} is not in {image}")
            image[dat_name].data = fft_lm_to_uv(image[sky_var].data, axes=[3, 4])


def ifft_to_lm(image, data_vars: Union[str, List[str]] = "") -> None:
    """standard_image_grid_preparation.py:344-377."""
    for vis_var in _vars(data_vars, ["VISIBILITY", "VISIBILITY_RESIDUAL", "VISIBILITY_MODEL"]):
        if vis_var in image.data_vars:
            dat_name = "SKY" if vis_var == "VISIBILITY" else vis_var.split("VISIBILITY_", 1)[1]
            image[dat_name].data = ifft_uv_to_lm(image[vis_var].data, axes=[3, 4])


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
