"""Convolution functions of the mosaic / A-projection gridder, on the GPU.

Mirror of core/imaging/make_gridding_convolution_function.py: the same
functions, arguments and results.  The per-pixel work runs on the device:
the airy-disk aperture patterns of the antenna types and the phase gradient
come from ``av_gcf_aperture_patterns`` and ``av_gcf_phase_gradient`` of
libavgrid.so, the baseline products, FFTs, crops and normalisations are
device tensor operations, and only the support search along the centre row
and column (a few hundred values) runs on the host.  The results are NumPy
arrays in a dataset with the reference's variable names and dimensions.

Two reference details are kept as they are:
* the antenna index of each baseline is taken as np.where(antenna_name ==
  baseline_antenna1_name)[0] with the comparison broadcast antenna-major, as
  xarray broadcasts it (:44-51);
* make_phase_gradient projects the field direction with astropy's
  RA---SIN/DEC--SIN WCS.  astropy is absent here, so the projection is
  restated (zenithal SIN, native longitude of the pole 180 degrees, unit PC
  matrix): the pixel offset from the reference pixel is (l / cell0, m /
  cell1) with l = cos d sin(a - a0), m = sin d cos d0 - cos d sin d0 cos(a -
  a0).
"""
from __future__ import annotations

import copy
import itertools

import numpy as np
import torch

from . import device as _dev
from ._lib import call
from .make_pb_symmetric import aperture_patterns_device, airy_disk_rorder, casa_airy_disk_rorder
from .msv4 import make_data_array

_GCF_DEFAULTS = {"function": "casa_airy", "chan_tolerance_factor": 0.005,
                 "oversampling": [10, 10], "max_support": [15, 15],
                 "support_cut_level": 2.5 * 10 ** -2, "a_chan_num_chunk": 3}


class GcfDataset(dict):
    """The convolution-function dataset: variables by name (also as
    attributes), ``attrs`` and ``coords``.  ``len`` counts the variables, as
    for an xarray Dataset."""

    def __init__(self):
        super().__init__()
        self.attrs = {}
        self.coords = {}

    def __getattr__(self, name):
        if name in self:
            return self[name]
        dims = {d: s for v in self.values() for d, s in zip(v.dims, v.data.shape)}
        if name in dims:
            return _Dim(dims[name])
        raise AttributeError(name)


class _Dim:
    def __init__(self, size):
        self.size = size


def _check_gcf_params(gcf_params):
    """Defaults and conversions of check_gcf_params
    (check_imaging_parameters.py:78-170)."""
    for k, v in _GCF_DEFAULTS.items():
        gcf_params.setdefault(k, v)
    if gcf_params["function"] not in ("casa_airy", "airy"):
        raise AssertionError("######### ERROR: gcf_params checking failed")
    if len(gcf_params["list_dish_diameters"]) != len(gcf_params["list_blockage_diameters"]):
        raise AssertionError("######### ERROR: gcf_params checking failed")
    for k in ("oversampling", "max_support"):
        gcf_params[k] = np.array(gcf_params[k]).astype(int)
    for k in ("freq_chan", "list_dish_diameters", "list_blockage_diameters", "unique_ant_indx",
              "pol"):
        gcf_params[k] = np.array(gcf_params[k])
    return gcf_params


def _data_group_in_name(sel_params):
    return (sel_params or {}).get("data_group_in_name", "base")


def make_gridding_convolution_function(gcf_xds, ms_xds, gcf_params, grid_params, sel_params):
    """make_gridding_convolution_function.py:24-193.

    ms_xds provides the attributes the reference reads from an xradio MSv4
    dataset: ``xr_ms.get_field_and_source_xds(name).FIELD_PHASE_CENTER_DIRECTION``,
    ``antenna_xds.antenna_name``, ``baseline_antenna1_name``,
    ``baseline_antenna2_name``, ``frequency`` and ``polarization``."""
    _gcf_params = copy.deepcopy(gcf_params)
    _grid_params = copy.deepcopy(grid_params)
    name = _data_group_in_name(sel_params)
    fpc = ms_xds.xr_ms.get_field_and_source_xds(name).FIELD_PHASE_CENTER_DIRECTION
    fpc = fpc.isel(field_name=0) if hasattr(fpc, "isel") else fpc
    _gcf_params["field_phase_dir"] = np.asarray(getattr(fpc, "values", fpc),
                                                dtype=np.float64).reshape(-1)[:2]
    names = np.asarray(getattr(ms_xds.antenna_xds.antenna_name, "values",
                               ms_xds.antenna_xds.antenna_name))
    b1 = np.asarray(getattr(ms_xds.baseline_antenna1_name, "values",
                            ms_xds.baseline_antenna1_name))
    b2 = np.asarray(getattr(ms_xds.baseline_antenna2_name, "values",
                            ms_xds.baseline_antenna2_name))
    # the antenna-major broadcast of (antenna_name == baseline_antennaN_name)
    a1 = np.where(names[:, None] == b1[None, :])[0]
    a2 = np.where(names[:, None] == b2[None, :])[0]
    _gcf_params["basline_ant"] = np.array([a1, a2]).T
    _gcf_params["freq_chan"] = np.asarray(getattr(ms_xds.frequency, "values", ms_xds.frequency))
    _gcf_params["pol"] = np.asarray(getattr(ms_xds.polarization, "values", ms_xds.polarization))
    _gcf_params = _check_gcf_params(_gcf_params)
    _gcf_params["resize_conv_size"] = (_gcf_params["max_support"] + 1) * _gcf_params[
        "oversampling"]

    nbl = len(_gcf_params["basline_ant"])
    if gcf_xds is None or len(gcf_xds) < 1 or gcf_xds.baseline.size != nbl:
        gcf_xds = GcfDataset()
        casa = _gcf_params["function"] == "casa_airy"
        n_unique_ant = len(_gcf_params["list_dish_diameters"])
        cf_baseline_map, pb_ant_pairs = create_cf_baseline_map(
            _gcf_params["unique_ant_indx"], _gcf_params["basline_ant"], n_unique_ant)
        cf_chan_map, pb_freq = create_cf_chan_map(_gcf_params["freq_chan"],
                                                  _gcf_params["chan_tolerance_factor"])
        cf_pol_map = np.zeros((len(_gcf_params["pol"]),), dtype=int)
        pb_pol = np.array([0])
        kernels = []
        for ipower in (1, 2):
            _gcf_params["ipower"] = ipower
            bp = _baseline_patterns_device(pb_freq, pb_ant_pairs, casa, _gcf_params,
                                           _grid_params)
            kernels.append(_fft_real(bp).unsqueeze(2))  # [pair, freq, pol = 1, u, v]
            del bp
        rk, rkc, support = resize_and_calc_support(kernels[0], kernels[1], _gcf_params,
                                                   _grid_params)
        gcf_xds["SUPPORT"] = make_data_array(support, None,
                                             ("conv_baseline", "conv_chan", "conv_pol", "xy"))
        gcf_xds["CONV_KERNEL_CONVOLVED"] = make_data_array(
            rkc, None, ("conv_baseline", "conv_chan", "conv_pol", "u", "v"))
        gcf_xds["CONV_KERNEL"] = make_data_array(
            rk, None, ("conv_baseline", "conv_chan", "conv_pol", "u", "v"))
        gcf_xds["PS_CORR_IMAGE"] = make_data_array(np.ones(_grid_params["image_size"]), None,
                                                   ("l", "m"))
        gcf_xds["CF_BASELINE_MAP"] = make_data_array(cf_baseline_map, None, ("baseline",))
        gcf_xds["CF_CHAN_MAP"] = make_data_array(cf_chan_map, None, ("chan",))
        gcf_xds["CF_POL_MAP"] = make_data_array(cf_pol_map, None, ("pol",))
        gcf_xds.attrs["cell_uv"] = 1 / (np.asarray(_grid_params["image_size_padded"]) *
                                        np.asarray(_grid_params["cell_size"]) *
                                        _gcf_params["oversampling"])
        gcf_xds.attrs["oversampling"] = _gcf_params["oversampling"]

    phase_gradient = make_phase_gradient(_gcf_params["field_phase_dir"][None, :], _gcf_params,
                                         _grid_params)
    gcf_xds["PHASE_GRADIENT"] = make_data_array(phase_gradient, None, ("field_id", "u", "v"))
    return gcf_xds


def _pb_grid_params(gcf_params, grid_params):
    # make_gridding_convolution_function.py:200-203
    pgp = copy.deepcopy(grid_params)
    pgp["cell_size"] = np.asarray(grid_params["cell_size"]) * gcf_params["oversampling"]
    pgp["image_size"] = np.asarray(pgp["image_size_padded"])
    pgp["image_center"] = pgp["image_size"] // 2
    return pgp


def _baseline_patterns_device(pb_freq, pb_ant_pairs, casa, gcf_params, grid_params):
    """[n_pair, n_freq, n0, n1] device tensor: the product of the two
    antenna types' patterns of each pair (:222-227)."""
    pat = aperture_patterns_device(pb_freq, gcf_params, _pb_grid_params(gcf_params, grid_params),
                                   casa)
    pairs = torch.as_tensor(np.asarray(pb_ant_pairs, dtype=np.int64).reshape(-1, 2),
                            device=pat.device)
    return pat[pairs[:, 0]] * pat[pairs[:, 1]]


def _fft_real(bp):
    """real(fftshift(fft2(ifftshift(bp)))) over the last two axes (:113-127)."""
    ax = (-2, -1)
    return torch.real(torch.fft.fftshift(torch.fft.fft2(torch.fft.ifftshift(bp, dim=ax), dim=ax),
                                         dim=ax)).contiguous()


def make_baseline_patterns(pb_freq, pb_pol, pb_ant_pairs, pb_func, gcf_params, grid_params):
    """make_gridding_convolution_function.py:196-229: [n_pair, n_freq, n_pol,
    n0, n1] float64; pb_func is airy_disk_rorder or casa_airy_disk_rorder."""
    if pb_func is airy_disk_rorder:
        casa = False
    elif pb_func is casa_airy_disk_rorder:
        casa = True
    else:
        raise ValueError("pb_func must be airy_disk_rorder or casa_airy_disk_rorder")
    bp = _baseline_patterns_device(pb_freq, pb_ant_pairs, casa, gcf_params, grid_params)
    out = np.zeros((bp.shape[0], bp.shape[1], len(pb_pol)) + tuple(bp.shape[2:]))
    out[:, :, 0] = bp.cpu().numpy()
    return out


def create_cf_baseline_map(unique_ant_indx, basline_ant, n_unique_ant):
    """make_gridding_convolution_function.py:232-257: the antenna-type pair
    of every baseline (pairs (i, j), i <= j, numbered row by row) and the
    pairs."""
    pairs = np.array([[i, j] for i in range(n_unique_ant) for j in range(i, n_unique_ant)],
                     dtype=int).reshape(-1, 2)
    types = np.asarray(unique_ant_indx)
    ba = np.asarray(basline_ant)
    t0, t1 = types[ba[:, 0]], types[ba[:, 1]]
    cf_baseline_map = np.zeros((ba.shape[0],), dtype=int)
    for k, (i, j) in enumerate(pairs):
        cf_baseline_map[(t0 == i) & (t1 == j)] = k
    return cf_baseline_map, pairs


def create_cf_chan_map(freq_chan, chan_tolerance_factor):
    """make_gridding_convolution_function.py:260-288: the primary-beam
    frequencies (one per tolerance-wide band, centred) and the nearest one
    of every channel; every channel its own when the bands outnumber them."""
    freq_chan = np.asarray(freq_chan)
    n_chan = len(freq_chan)
    tol = np.max(freq_chan) * chan_tolerance_factor
    n_pb_chan = int(np.floor((np.max(freq_chan) - np.min(freq_chan)) / tol) + 0.5)
    if n_pb_chan == 0:
        n_pb_chan = 1
    if n_pb_chan >= n_chan:
        return np.arange(n_chan), freq_chan
    width = (np.max(freq_chan) - np.min(freq_chan)) / n_pb_chan
    pb_freq = np.arange(n_pb_chan) * width + np.min(freq_chan) + width / 2
    cf_chan_map = np.zeros((n_chan,), dtype=int)
    for i in range(n_chan):
        cf_chan_map[i], _ = find_nearest(pb_freq, freq_chan[i])
    return cf_chan_map, pb_freq


def find_nearest(array, value):
    """make_gridding_convolution_function.py:291-294."""
    array = np.asarray(array)
    idx = (np.abs(array - value)).argmin()
    return idx, array[idx]


def _as_device(a):
    if isinstance(a, torch.Tensor):
        return a.to(_dev.require_gpu(), torch.float64)
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(_dev.require_gpu())


def resize_and_calc_support(conv_kernel, conv_kernel_convolved, gcf_params, grid_params):
    """make_gridding_convolution_function.py:297-371: the support of every
    (baseline pair, channel, pol) kernel from its convolved kernel, the
    central resize_conv_size crop of both kernels, each normalised so that
    its sum over the support region is oversampling[0] * oversampling[1]."""
    ck, ckc = _as_device(conv_kernel), _as_device(conv_kernel_convolved)
    if ck.dim() != 5 or ckc.shape != ck.shape:
        raise ValueError("conv kernels must be [conv_baseline, conv_chan, conv_pol, u, v]")
    shape = tuple(ck.shape[:3])
    support = np.zeros(shape + (2,), dtype=int)
    rcs = np.asarray(gcf_params["resize_conv_size"]).astype(int)
    ov = np.asarray(gcf_params["oversampling"]).astype(int)
    npad = np.asarray(grid_params["image_size_padded"]).astype(int)
    s0 = npad // 2 - rcs // 2
    e0 = s0 + rcs
    rk = torch.empty(shape + tuple(rcs), dtype=torch.float64, device=ck.device)
    rkc = torch.empty_like(rk)
    for idx in itertools.product(*[range(s) for s in shape]):
        support[idx] = calc_conv_size(ckc[idx], npad, gcf_params["support_cut_level"], ov,
                                      gcf_params["max_support"])
        emb = (support[idx] + 1) * ov
        es = rcs // 2 - emb // 2
        ee = es + emb
        for src, dst in ((ck, rk), (ckc, rkc)):
            crop = src[idx][s0[0]:e0[0], s0[1]:e0[1]]
            norm = torch.sum(crop[es[0]:ee[0], es[1]:ee[1]]) / (ov[0] * ov[1])
            dst[idx] = crop / norm
    return rk.cpu().numpy(), rkc.cpu().numpy(), support


def calc_conv_size(sub_a_term, imsize, support_cut_level, oversampling, max_support):
    """make_gridding_convolution_function.py:374-423: [support_x, support_y]
    (odd, equal) from where the kernel falls to support_cut_level of its
    peak along the centre row and column."""
    a = _as_device(sub_a_term)
    absa = torch.abs(a)
    min_amplitude, max_amplitude = float(absa.min()), float(absa.max())
    cut = support_cut_level * max_amplitude
    assert min_amplitude < cut, "######### ERROR: support_cut_level too small or imsize too small."
    cx, cy = int(imsize[0]) // 2, int(imsize[1]) // 2
    col = a[cx:, cy].cpu().numpy()  # sub_a_term[indx_x, cy] for indx_x >= cx
    row = a[cx, cy:].cpu().numpy()
    supports = []
    for line, n, ov in ((col, imsize[0], oversampling[0]), (row, imsize[1], oversampling[1])):
        k = 0
        while line[k] > cut:
            k += 1
            assert k < n - n // 2, \
                "######### ERROR: support_cut_level too small or imsize too small."
        supports.append((int(0.5 + k / ov) + 1) * 2 + 1)
    support_x, support_y = supports
    assert support_x < max_support[0], (
        "######### ERROR: support_cut_level too small or imsize too small." + str(support_x) +
        ",*," + str(max_support[0]))
    assert support_y < max_support[1], (
        "######### ERROR: support_cut_level too small or imsize too small." + str(support_y) +
        ",*," + str(max_support[1]))
    if support_x > support_y:
        support_y = support_x
    else:
        support_x = support_y
    return [support_x, support_y]


def sin_pixel_offset(field_dir, phase_center, cell):
    """Offset in pixels of directions field_dir [n, 2] (RA, Dec in radians)
    from the reference pixel of a RA---SIN/DEC--SIN image centred on
    phase_center with pixel steps cell (radians)."""
    fd = np.atleast_2d(np.asarray(field_dir, dtype=np.float64))
    a, d = fd[:, 0], fd[:, 1]
    a0, d0 = float(phase_center[0]), float(phase_center[1])
    l = np.cos(d) * np.sin(a - a0)
    m = np.sin(d) * np.cos(d0) - np.cos(d) * np.sin(d0) * np.cos(a - a0)
    return np.stack([l / cell[0], m / cell[1]], axis=1)


def make_phase_gradient(field_phase_dir, gcf_params, grid_params):
    """make_gridding_convolution_function.py:426-485: [n_field, resize0,
    resize1] complex128 phase ramps that shift each field's aperture to the
    image phase centre."""
    phase_center = np.asarray(getattr(gcf_params["phase_direction"], "values",
                                      gcf_params["phase_direction"]), dtype=np.float64)
    npad = np.asarray(grid_params["image_size_padded"])
    pix_dist = sin_pixel_offset(field_phase_dir, phase_center,
                                np.asarray(grid_params["cell_size"], dtype=np.float64))
    pix = -(pix_dist) * 2 * np.pi / (npad * np.asarray(gcf_params["oversampling"]))
    size = np.asarray(gcf_params["resize_conv_size"]).astype(int)
    centre = size // 2
    dev = _dev.require_gpu()
    pix_t = torch.from_numpy(np.ascontiguousarray(pix, dtype=np.float64)).to(dev)
    out = torch.empty((pix.shape[0], int(size[0]), int(size[1])), dtype=torch.complex128,
                      device=dev)
    call("av_gcf_phase_gradient", _dev._ptr(pix_t), pix.shape[0], int(size[0]), int(size[1]),
         int(centre[0]), int(centre[1]), _dev._ptr(out), _dev._stream())
    return out.cpu().numpy()
