"""calculate_imaging_weights on the GPU (mirror of
core/imaging/calculate_imaging_weights.py:28-275) for processing sets given
as a mapping of names to MSv4-shaped datasets (or a list of them).

Natural weighting leaves the data weights as the imaging weights.  Briggs
weighting grids every dataset's data weights (flagged samples NaN; two
polarisations averaged, four: the first and last averaged) into one
float64 weight-density grid per channel, forms the Briggs factors and
writes WEIGHT_IMAGING = data weight / (a * G + b), the same for every
polarisation.  The data-group bookkeeping of xradio is kept as a plain
dictionary in each dataset's attrs["data_groups"].
"""
from __future__ import annotations

import copy
from datetime import datetime, timezone
from typing import Dict

import numpy as np

from .imaging_weighting.briggs_weighting import calculate_briggs_params
from .imaging_weighting.grid_imaging_weights import (degrid_imaging_weights,
                                                     grid_imaging_weights)
from .msv4 import make_data_array


def _datasets(ps_xdt):
    return list(ps_xdt.values()) if hasattr(ps_xdt, "values") and not hasattr(
        ps_xdt, "data_vars") else list(ps_xdt)


def _single_pol(ms, names):
    w = np.array(ms[names["weight"]].data, dtype=np.float64, copy=True)
    w[np.asarray(ms[names["flag"]].data) == 1] = np.nan
    if w.shape[3] == 2:
        w = ((w[..., 0] + w[..., 1]) / 2)[..., np.newaxis]
    if w.shape[3] == 4:
        w = ((w[..., 0] + w[..., 3]) / 2)[..., np.newaxis]
    return w


def calculate_imaging_weights(ps_xdt, grid_params: Dict, imaging_weights_params: Dict,
                              sel_params: Dict, return_weight_density_grid: bool = False):
    """Returns (ps_xdt, data_group_out) or, with return_weight_density_grid,
    (ps_xdt, data_group_out, weight_density_grid (n_chan, 1, n_u, n_v))."""
    iw = copy.deepcopy(imaging_weights_params)
    if iw.get("weighting") not in ("natural", "briggs", "briggs_abs"):
        raise AssertionError("######### ERROR: imaging_weights_params checking failed")
    sel = copy.deepcopy(sel_params)
    out_name = sel.get("data_group_out_name", "imaging")
    natural = iw["weighting"] == "natural"
    names = {"weight": "WEIGHT", "uvw": "UVW", "flag": "FLAG",
             "weight_imaging": "WEIGHT" if natural else "WEIGHT_IMAGING"}
    names.update(sel.get("data_group_out", {}))
    group = dict(names)
    dss = _datasets(ps_xdt)
    stamp = datetime.now(timezone.utc).isoformat()
    if natural:
        for ms in dss:
            ms.attrs.setdefault("data_groups", {})[out_name] = dict(
                group, date=stamp, description="Data group created for natural imaging weights "
                                               "with .")
        return ps_xdt, dict(group, data_group_out_name=out_name)
    n_uv = grid_params["image_size"]
    gp = {"image_size": n_uv, "cell_size": grid_params["cell_size"]}
    freqs = [np.asarray(ms.coords["frequency"].values) for ms in dss]
    n_chan = freqs[0].size
    wgrid = np.zeros((n_chan, 1, int(n_uv[0]), int(n_uv[1])), dtype=np.double)
    sum_weight = np.zeros((n_chan, 1), dtype=np.double)
    for ms, f in zip(dss, freqs):
        grid_imaging_weights(wgrid, sum_weight, np.asarray(ms[names["uvw"]].data),
                             _single_pol(ms, names), f, gp)
    bf = calculate_briggs_params(wgrid, sum_weight, iw)
    desc = ("Data group created for briggs imaging weights with robust value "
            + str(iw.get("robust")))
    for ms, f in zip(dss, freqs):
        wi = degrid_imaging_weights(wgrid, np.asarray(ms[names["uvw"]].data),
                                    _single_pol(ms, names), bf, f, gp)
        n_pol = ms.sizes["polarization"]
        ref = ms[names["weight"]]
        ms[names["weight_imaging"]] = make_data_array(np.tile(wi, (1, 1, 1, n_pol)),
                                                      coords=ref.coords, dims=ref.dims)
        ms.attrs.setdefault("data_groups", {})[out_name] = dict(group, date=stamp,
                                                                description=desc)
    out = dict(group, data_group_out_name=out_name)
    if return_weight_density_grid:
        return ps_xdt, out, wgrid
    return ps_xdt, out
