"""The reference's own hot-path tests, run through the GPU mirror of its API.

Each test restates a test of tests/unit/core/imaging/... on the same kind of
input (synthetic uvw where the reference downloads a measurement set) and
also compares with the FP64 oracle.
"""
import json
import os

import numpy as np
import pytest

pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

C = 299792458.0
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")


def rel(a, b):
    den = np.max(np.abs(b))
    return float(np.max(np.abs(np.asarray(a) - b)) / (den if den > 0 else 1.0))


def gold(name):
    with open(GOLDEN) as f:
        return json.load(f)[name]


def point_source_setup():
    """test_standard_grid.py:15-41: 200 x 200 regular uv grid carrying
    conj(FFT(model)) of two point sources."""
    from astroviper_amd.fft import fft_lm_to_uv
    g = gold("standard_grid_point_sources")
    nant, maxuv = g["nant"], g["maxUV"]
    freq = np.array([g["freq"]])
    cell = 0.5 * C / freq[0] / maxuv
    mod = np.zeros((nant, nant))
    for (x, y, s) in g["sources"]:
        mod[x, y] = s
    ft = np.conj(fft_lm_to_uv(mod, axes=[0, 1]))
    ax = np.linspace(-maxuv, maxuv, nant)
    vis = np.zeros((1, nant * nant, 1, 1), dtype=np.complex128)
    uvw = np.zeros((1, nant * nant, 3))
    for ui, u in enumerate(ax):
        for vi, v in enumerate(ax):
            b = ui * nant + vi
            uvw[0, b] = [u, v, 0.0]
            vis[0, b, 0, 0] = ft[ui, vi]
    return g, vis, uvw, np.ones(vis.shape), freq, np.array([nant, nant]), np.array([cell, cell])


def test_standard_grid_numpy_wrap_input_checked(gpu, oracle):
    from astroviper_amd.gcf_prolate_spheroidal import (create_prolate_spheroidal_kernel,
                                                       create_prolate_spheroidal_kernel_1D)
    from astroviper_amd.ifft import ifft_uv_to_lm
    from astroviper_amd.standard_grid import standard_grid_numpy_wrap_input_checked
    g, vis, uvw, w, freq, gs, cs = point_source_setup()
    cgk = create_prolate_spheroidal_kernel_1D(100, 7)
    grid, sumwt = standard_grid_numpy_wrap_input_checked(
        vis_data=vis, grid=np.array([]), sum_weight=np.array([]), uvw=uvw, weight=w,
        freq_chan=freq, cgk_1D=cgk, image_size=gs, cell_size=cs, oversampling=100, support=7,
        complex_grid=True, do_psf=False, chan_mode="continuum")
    assert grid.dtype == np.complex128 and grid.shape == (1, 1, 200, 200)
    _, corr = create_prolate_spheroidal_kernel(100, 7, gs)
    dirty = ifft_uv_to_lm(grid, axes=[2, 3]) / corr * gs[0] * gs[1] / sumwt
    for (x, y, thr) in g["dirty_thresholds"]:
        assert dirty[0, 0, x, y] > thr
    gref, swref = oracle.standard_grid(vis, uvw, w, freq, gs, cs)
    assert np.max(np.abs(grid - gref)) <= 1e-6 * np.max(np.abs(gref))
    grid, sumwt = standard_grid_numpy_wrap_input_checked(
        vis_data=vis, grid=np.array([]), sum_weight=np.array([]), uvw=uvw, weight=w,
        freq_chan=freq, cgk_1D=cgk, image_size=gs, cell_size=cs, do_psf=True)
    psf = ifft_uv_to_lm(grid, axes=[2, 3]) / corr * gs[0] * gs[1] / sumwt
    assert abs(psf[0, 0, 100, 100] - 1.0) < g["psf_centre_tol"]
    # accumulation into a supplied grid doubles it (standard_grid.py:133-143)
    g2, s2 = standard_grid_numpy_wrap_input_checked(
        vis_data=vis, grid=grid.copy(), sum_weight=sumwt.copy(), uvw=uvw, weight=w,
        freq_chan=freq, cgk_1D=cgk, image_size=gs, cell_size=cs, do_psf=True)
    np.testing.assert_allclose(g2, 2 * grid, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(s2, 2 * sumwt, rtol=1e-12)


def cube_dataset(nchan=16, npol=2, n=200, seed=2):
    """A measurement set whose channel k carries the constant k + 1
    (standard_gridding_example.py:23-52 on synthetic uvw)."""
    from astroviper_amd import synth
    from astroviper_amd.msv4 import make_visibility_dataset
    arr = synth.ArrayConfig(n_ant=12, n_time=30, n_chan=nchan, n_pol=npol, seed=seed,
                            freq_lo=1.0e9, freq_hi=1.05e9)
    obs = synth.make_observation(arr, synth.SkyConfig())
    maxuv = np.max(np.hypot(obs.uvw[..., 0], obs.uvw[..., 1]))
    cell = 0.5 * C / obs.freq[-1] / maxuv
    nt, nb = obs.uvw.shape[:2]
    vis = np.ones((nt, nb, nchan, npol), complex) * (np.arange(nchan) + 1)[None, None, :, None]
    ds = make_visibility_dataset(vis, obs.uvw.copy(), np.ones(vis.shape), np.zeros(vis.shape, bool),
                                 obs.freq)
    return cell, ds


def test_grid2image_spheroid_ms4_cube_constant(gpu, oracle):
    from astroviper_amd.standard_grid import grid2image_spheroid_ms4
    cell, ms4 = cube_dataset()
    dirty = np.zeros([16, 2, 200, 200])
    grid2image_spheroid_ms4(vis=ms4, resid_array=dirty, pixelincr=np.array([-cell, cell]),
                            support=7, sampling=100, dopsf=False, column="VISIBILITY",
                            chan_mode="cube")
    for k in range(16):
        # test_standard_grid.py:126-132
        assert round(dirty[k, 0, 100, 100] - np.round(float(k + 1)), 4) == 0
    ref, _, _ = oracle.grid2image(ms4.VISIBILITY.data, ms4.UVW.data, ms4.WEIGHT.data,
                                  ms4.FLAG.data, ms4.coords["frequency"].values, (200, 200),
                                  (-cell, cell), chan_mode="cube")
    assert np.max(np.abs(dirty - ref)) <= 1e-6 * np.max(np.abs(ref))


def test_grid2image_mutates_weights_and_checks_shapes(gpu):
    from astroviper_amd.standard_grid import grid2image_spheroid_ms4
    cell, ms4 = cube_dataset(nchan=4)
    ms4.FLAG.data[0, :3] = True
    ms4.UVW.data[1, 2, 0] = np.nan
    dirty = np.zeros([1, 2, 200, 200])
    grid2image_spheroid_ms4(ms4, dirty, np.array([-cell, cell]))
    assert np.all(ms4.WEIGHT.data[0, :3] == 0) and np.all(ms4.WEIGHT.data[1, 2] == 0)
    with pytest.raises(RuntimeError, match="Shapes of gridded vis and image do no match"):
        grid2image_spheroid_ms4(ms4, np.zeros([2, 2, 200, 200]), np.array([-cell, cell]))
    with pytest.raises(TypeError, match="not an xarray datatree"):
        grid2image_spheroid_ms4([ms4, "not a dataset"], dirty, np.array([-cell, cell]))


def test_grid2image_list_of_datasets_accumulates(gpu, oracle):
    from astroviper_amd.standard_grid import grid2image_spheroid_ms4
    cell, a = cube_dataset(nchan=3, seed=4)
    _, b = cube_dataset(nchan=3, seed=5)
    both = np.zeros([1, 2, 200, 200])
    grid2image_spheroid_ms4([a, b], both, np.array([-cell, cell]))
    uvw = np.concatenate([a.UVW.data, b.UVW.data], axis=1)
    vis = np.concatenate([a.VISIBILITY.data, b.VISIBILITY.data], axis=1)
    w = np.ones(vis.shape)
    ref, _, _ = oracle.grid2image(vis, uvw, w, None, a.coords["frequency"].values, (200, 200),
                                  (-cell, cell))
    assert np.max(np.abs(both - ref)) <= 1e-6 * np.max(np.abs(ref))


def test_dgrid_numba_reference_call(gpu, oracle):
    """test_standard_degrid.py:183-210 (dgrid_numba with scale = cell_size,
    offset = grid_size // 2) on a constant grid of 2 -> max |vis| == 2.0."""
    from astroviper_amd.gcf_prolate_spheroidal import create_prolate_spheroidal_kernel_1D
    from astroviper_amd.standard_degrid import dgrid_numba
    g = gold("dgrid_constant_grid")
    npix, maxuv = g["npix"], g["maxUV"]
    freq = np.array([g["freq"]])
    cell = 0.5 * C / freq[0] / maxuv
    rng = np.random.default_rng(9)
    uvw = np.zeros((1, g["n_uvw"], 3))
    uvw[0, :, :2] = rng.uniform(-maxuv, maxuv, size=(g["n_uvw"], 2))
    grid = np.full((1, 1, npix, npix), g["grid_value"] + 0j)
    vals = np.zeros((1, g["n_uvw"], 1, 1), dtype=np.complex128)
    flag = np.zeros(vals.shape, bool)
    conv = create_prolate_spheroidal_kernel_1D(100, 7)
    dgrid_numba(uvw, np.zeros((1, g["n_uvw"])), vals, flag, np.array([cell, cell]),
                np.array([npix, npix]) // 2, grid, freq, C, 7, 100, conv,
                np.zeros(1, dtype=int), np.zeros(1, dtype=int))
    assert np.max(np.abs(vals)) == g["expected_max_abs"]
    # against the oracle on a random grid, same call
    rg = rng.standard_normal((1, 1, npix, npix)) + 1j * rng.standard_normal((1, 1, npix, npix))
    v1 = np.zeros_like(vals)
    dgrid_numba(uvw, np.zeros((1, g["n_uvw"])), v1, flag, np.array([cell, cell]),
                np.array([npix, npix]) // 2, rg, freq, C, 7, 100, conv,
                np.zeros(1, dtype=int), np.zeros(1, dtype=int))
    v2 = np.zeros_like(vals)
    oracle.dgrid(uvw, v2, flag, np.array([cell, cell, 0]), np.array([npix // 2, npix // 2, 0]), rg,
                 freq, 7, 100, conv, np.zeros(1, dtype=int), np.zeros(1, dtype=int))
    assert np.max(np.abs(v1 - v2)) <= 1e-6 * np.max(np.abs(v2))


def test_degrid_spheroid_ms4(gpu, oracle):
    """test_standard_degrid.py:242-257 plus parity and the NaN-uvw mutation."""
    from astroviper_amd.standard_degrid import degrid_spheroid_ms4
    cell, ms = cube_dataset(nchan=3)
    ms.UVW.data[2, 5, 1] = np.nan
    ms.FLAG.data[4, :2, 1, 0] = True
    const = np.full((1, 1, 200, 200), 2.0 + 0j)
    degrid_spheroid_ms4(vis=ms, grid=const, pixelincr=np.array([cell, cell]), support=7,
                        sampling=100, whichFunc=0)
    m = ms.VISIBILITY_MODEL.data
    assert np.max(m) == 2.0 and m.dtype == np.complex128
    assert ms.VISIBILITY_MODEL.dims == ms.VISIBILITY.dims
    assert np.all(ms.FLAG.data[2, 5]) and np.all(ms.UVW.data[2, 5] == 1e9)
    assert np.all(m[2, 5] == 0) and np.all(m[4, :2, 1, 0] == 0)
    rng = np.random.default_rng(10)
    grid = rng.standard_normal((3, 2, 200, 200)) + 1j * rng.standard_normal((3, 2, 200, 200))
    degrid_spheroid_ms4(ms, grid, np.array([cell, cell]))
    ref = oracle.degrid(ms.UVW.data, ms.FLAG.data, ms.coords["frequency"].values, grid,
                        (cell, cell))
    assert np.max(np.abs(ms.VISIBILITY_MODEL.data - ref)) <= 1e-6 * np.max(np.abs(ref))
    # incremental adds to the existing model (standard_degrid.py:540-551)
    before = ms.VISIBILITY_MODEL.data.copy()
    degrid_spheroid_ms4(ms, grid, np.array([cell, cell]), incremental=True)
    np.testing.assert_allclose(ms.VISIBILITY_MODEL.data, 2 * before, rtol=1e-12, atol=1e-14)


def test_make_psf_and_imager_calls(gpu, oracle):
    from astroviper_amd.imager import (compute_residual_visibilities, grid_visibilities,
                                       predict_model_visibilities)
    from astroviper_amd.make_point_spread_function import make_psf
    cell, ms = cube_dataset(nchan=2)
    psf = make_psf(ms, {"cell_size": (cell, cell), "image_size": (200, 200),
                        "phase_center": (0.0, 0.5), "chan_mode": "cube"},
                   {"sampling": 100, "support": 7})
    p = psf["POINT_SPREAD_FUNCTION"].data
    assert p.shape == (1, 2, 2, 200, 200)
    assert psf["POINT_SPREAD_FUNCTION"].dims == ("time", "frequency", "polarization", "l", "m")
    pk = np.unravel_index(np.argmax(p[0, 0, 0]), (200, 200))
    assert abs(pk[0] - 100) <= 1 and abs(pk[1] - 100) <= 1 and p.max() > 0.1
    dirty = grid_visibilities(ms, (200, 200), (-cell, cell), chan_mode="cube")
    assert dirty.shape == (2, 2, 200, 200)
    model = np.zeros((2, 2, 200, 200))
    model[:, :, 100, 100] = 1.0
    predict_model_visibilities(ms, model, (cell, cell))
    compute_residual_visibilities(ms)
    assert ms.RESIDUAL.data.shape == ms.VISIBILITY.data.shape


def test_fft_round_trip_and_known_answers(gpu):
    """test_fft_ifft.py:9-24 and test_standard_image_grid_preparation.py:127-142."""
    from astroviper_amd.fft import fft_lm_to_uv
    from astroviper_amd.ifft import ifft_uv_to_lm
    g = gold("fft")
    sky = np.zeros((128, 128))
    sky[64, 64] = 1
    sky[50, 40] = 1
    rt = ifft_uv_to_lm(fft_lm_to_uv(sky, (0, 1)), (0, 1))
    assert np.max(np.abs(sky - rt)) < g["roundtrip_tol"]
    uv = fft_lm_to_uv(np.ones((200, 200)), (0, 1))
    assert abs(uv[100, 100] - g["ones_200_dc"]) < 1e-9
    vis = np.zeros((200, 200), complex)
    vis[100, 100] = 10
    assert abs(ifft_uv_to_lm(vis, (0, 1)).max() - g["delta_10_max"]) < 1e-15


# ------------------------------------ xradio-image gridding (A12 mirror)
def lm_image(nx=200, ny=200, nchan=1, npol=1, ntime=1, which="RESIDUAL", cell=4.8e-6,
             value=1.0):
    """test_standard_image_grid_preparation.py:17-65: a sky image of ones
    over (time, frequency, polarization, l, m), 4.8e-6 rad cells."""
    from astroviper_amd.msv4 import DataArray, ImageDataset
    coords = {"time": 5.7e4 + np.arange(ntime), "frequency": np.arange(1, nchan + 1) * 1e8 + 1e9,
              "polarization": np.array(["RR", "LL", "LR", "RL"])[:npol],
              "l": (np.arange(nx) - nx // 2) * -cell, "m": (np.arange(ny) - ny // 2) * cell}
    im = ImageDataset({}, coords)
    dims = ("time", "frequency", "polarization", "l", "m")
    im[which] = DataArray(np.full((ntime, nchan, npol, nx, ny), value, dtype=np.float32),
                          dims=dims, coords={k: coords[k] for k in dims})
    return im


def test_fft_to_uv_known_answers(gpu):
    """test_standard_image_grid_preparation.py:127-135: the FFT of a 200 x
    200 image of ones has DC 200*200, and 1.0 after correct_fft_to_uv; the
    pixel off the origin stays 0."""
    from astroviper_amd.standard_image_grid_preparation import (correct_fft_to_uv, fft_to_uv,
                                                                make_empty_padded_uv_image)
    a = lm_image()
    b = make_empty_padded_uv_image(a, image_size=[a.sizes["l"], a.sizes["m"]])
    assert b["VISIBILITY_RESIDUAL"].shape == (1, 1, 1, 200, 200)
    fft_to_uv(b)
    assert b["VISIBILITY_RESIDUAL"].sel(u=0, v=0).values[0] == 200 * 200
    correct_fft_to_uv(b)
    assert b["VISIBILITY_RESIDUAL"].sel(u=0, v=0).values[0] == 1.0
    assert b["VISIBILITY_RESIDUAL"].isel(u=99, v=99).values[0] == 0.0
    # padding keeps the image centred and the DC term
    c = make_empty_padded_uv_image(a, image_size=[256, 256])
    assert c["RESIDUAL"].shape == (1, 1, 1, 256, 256)
    assert c["RESIDUAL"].data[0, 0, 0, 28:228, 28:228].min() == 1.0
    assert c["RESIDUAL"].data.sum() == 200 * 200
    fft_to_uv(c)
    assert c["VISIBILITY_RESIDUAL"].sel(u=0, v=0).values[0] == 200 * 200
    with pytest.raises(Exception, match="larger than requested"):
        make_empty_padded_uv_image(a, image_size=[100, 100])


def test_ifft_to_lm_known_answer(gpu):
    """test_standard_image_grid_preparation.py:137-142: a delta of 10 at the
    uv origin transforms to a flat image of 10/200/200."""
    from astroviper_amd.standard_image_grid_preparation import (ifft_to_lm,
                                                                make_empty_padded_uv_image)
    a = lm_image()
    b = make_empty_padded_uv_image(a, image_size=[a.sizes["l"], a.sizes["m"]])
    b["VISIBILITY_RESIDUAL"].sel(u=0, v=0).values[0] = 10
    ifft_to_lm(b)
    r = b["RESIDUAL"].data
    assert abs(r.max() - 10.0 / 200.0 / 200.0) < 1e-10
    assert abs(r.min() - 10.0 / 200.0 / 200.0) < 1e-10
    # the result replaces the variable's data (a real image)
    assert not np.iscomplexobj(r)


def test_fft_to_uv_replaces_data(gpu):
    """fft_to_uv assigns the transform to the variable (the reference's
    `.data = `), so a complex result is kept even where the target held
    real values, and the variable takes the transform's dtype."""
    from astroviper_amd.msv4 import DataArray
    from astroviper_amd.standard_image_grid_preparation import fft_to_uv
    a = lm_image(nx=64, ny=64, which="SKY", value=0.0)
    a["SKY"].data[0, 0, 0, 10, 20] = 1.0
    dims = ("time", "frequency", "polarization", "l", "m")
    a["VISIBILITY_SKY"] = DataArray(np.zeros((1, 1, 1, 64, 64)), dims=dims)
    fft_to_uv(a, "SKY")
    v = a["VISIBILITY_SKY"].data
    assert np.iscomplexobj(v)
    ref = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(a["SKY"].data.astype(np.float64),
                                                       axes=[3, 4]), axes=[3, 4]), axes=[3, 4])
    assert np.max(np.abs(v - ref)) < 1e-5
    assert np.abs(v.imag).max() > 0.5


def uv_image_for(ms, n, cell, ntime=2, data_var="VISIBILITY"):
    """An image dataset with l/m and u/v coordinates for gridding `ms`
    (cube: one plane per channel) at `n` x `n`, `ntime` time planes."""
    from astroviper_amd.msv4 import DataArray, ImageDataset
    nchan = len(ms.coords["frequency"].values)
    npol = len(ms.coords["polarization"].values)
    coords = {"time": np.arange(ntime, dtype=float), "frequency": ms.coords["frequency"].values,
              "polarization": ms.coords["polarization"].values,
              "l": (np.arange(n) - n // 2) * -cell, "m": (np.arange(n) - n // 2) * cell,
              "u": (np.arange(n) - n // 2) / (n * cell), "v": (np.arange(n) - n // 2) / (n * cell)}
    im = ImageDataset({}, coords)
    uvd = ("time", "frequency", "polarization", "u", "v")
    skd = ("time", "frequency", "polarization", "l", "m")
    im[data_var] = DataArray(np.zeros((ntime, nchan, npol, n, n), np.complex64), dims=uvd)
    im[data_var + "_NORMALIZATION"] = DataArray(np.zeros((ntime, nchan, npol)), dims=uvd[:3])
    im["SKY"] = DataArray(np.zeros((ntime, nchan, npol, n, n), np.float32), dims=skd)
    return im


def test_grid2xradio_matches_oracle(gpu, oracle):
    """grid2xradio_spheroid_ms4 (standard_image_grid_preparation.py:495-580)
    grids every dataset of the list into every time plane of the complex64
    uv image and its normalisation; ifft_to_lm then correct_ifft_to_lm with
    doSpheroidCorr (:412-492) give the dirty image of
    grid2image_spheroid_ms4.  Compared with the FP64 oracle at the FP32
    bars (the uv image is complex64, so gridding and the FFT run in FP32)."""
    from astroviper_amd.standard_image_grid_preparation import (correct_ifft_to_lm,
                                                                grid2xradio_spheroid_ms4,
                                                                ifft_to_lm)
    cell, a = cube_dataset(nchan=3, seed=4)
    _, b = cube_dataset(nchan=3, seed=5)
    rng = np.random.default_rng(13)
    for ds in (a, b):
        ds.VISIBILITY.data[...] = (rng.standard_normal(ds.VISIBILITY.data.shape)
                                   + 1j * rng.standard_normal(ds.VISIBILITY.data.shape))
        ds.WEIGHT.data[...] = rng.uniform(0.5, 1.5, ds.WEIGHT.data.shape)
        ds.FLAG.data[...] = rng.random(ds.FLAG.data.shape) < 0.1
    uvw = np.concatenate([a.UVW.data, b.UVW.data], axis=1)
    vis = np.concatenate([a.VISIBILITY.data, b.VISIBILITY.data], axis=1)
    w = np.concatenate([a.WEIGHT.data, b.WEIGHT.data], axis=1)
    flag = np.concatenate([a.FLAG.data, b.FLAG.data], axis=1)
    freq = a.coords["frequency"].values
    n = 200
    w0 = np.where(flag, 0.0, w)
    gref, swref = oracle.standard_grid(vis, uvw, w0, freq, (n, n), (cell, cell),
                                       chan_mode="cube")
    im = uv_image_for(a, n, cell)
    grid2xradio_spheroid_ms4([a, b], im, data_var="VISIBILITY", chan_mode="cube")
    # flagged weights are zeroed in the caller's datasets, as in the reference
    assert np.all(a.WEIGHT.data[a.FLAG.data] == 0)
    # a complex64 uv image grids in FP32, so the weights are rounded to
    # float32 on the way in (the sums of weights are still formed in FP64)
    for t in range(2):
        e = rel(im["VISIBILITY"].data[t], gref)
        assert e < 1e-4, (t, e)
        np.testing.assert_allclose(im["VISIBILITY_NORMALIZATION"].data[t], swref, rtol=1e-6)
    ifft_to_lm(im, "VISIBILITY")
    correct_ifft_to_lm(im, "SKY", doSpheroidCorr=True)
    iref = oracle.normalise(gref, swref)
    _, corr = oracle.create_prolate_spheroidal_kernel(100, 7, np.array([n, n]))
    inner = corr >= 1e-2
    for t in range(2):
        d = np.abs(im["SKY"].data[t] - iref)
        full, inn = d.max() / np.abs(iref).max(), d[..., inner].max() / np.abs(iref).max()
        print(f"grid2xradio dirty plane {t}: full field {full:.2e}, taper >= 1e-2 {inn:.2e}")
        assert inn < 1e-4 and full < 1e-2
    # UV_SAMPLING grids the weights (the PSF grid)
    pim = uv_image_for(a, n, cell, ntime=1, data_var="UV_SAMPLING")
    grid2xradio_spheroid_ms4([a, b], pim, data_var="UV_SAMPLING", chan_mode="cube")
    pref, pswref = oracle.standard_grid(None, uvw, w0, freq, (n, n), (cell, cell),
                                        chan_mode="cube", do_psf=True)
    assert rel(pim["UV_SAMPLING"].data[0], pref) < 1e-4
    np.testing.assert_allclose(pim["UV_SAMPLING_NORMALIZATION"].data[0], pswref, rtol=1e-6)
    # continuum into a one-channel image, and the errors of the reference
    cim = uv_image_for(a, n, cell, ntime=1)
    cim["VISIBILITY"].data = np.zeros((1, 1, 2, n, n), np.complex64)
    cim["VISIBILITY_NORMALIZATION"].data = np.zeros((1, 1, 2))
    grid2xradio_spheroid_ms4(a, cim, data_var="VISIBILITY", chan_mode="continuum")
    cref, _ = oracle.standard_grid(a.VISIBILITY.data, a.UVW.data, a.WEIGHT.data, freq, (n, n),
                                   (cell, cell), chan_mode="continuum")
    assert rel(cim["VISIBILITY"].data[0], cref) < 1e-4
    with pytest.raises(Exception, match="VISIBILITY_MODEL is not in"):
        grid2xradio_spheroid_ms4(a, cim, data_var="VISIBILITY_MODEL")
    with pytest.raises(TypeError, match="not an xarray datatree"):
        grid2xradio_spheroid_ms4([a, "x"], cim)
    nosw = uv_image_for(a, n, cell, ntime=1)
    del nosw.data_vars["VISIBILITY_NORMALIZATION"]
    with pytest.raises(Exception, match="Do not have VISIBILITY_NORMALIZATION"):
        correct_ifft_to_lm(nosw, "SKY", doSpheroidCorr=True)


def test_correct_ifft_to_lm_plain_scaling(gpu):
    """correct_ifft_to_lm without doSpheroidCorr multiplies by nx*ny."""
    from astroviper_amd.standard_image_grid_preparation import correct_ifft_to_lm
    a = lm_image(nx=32, ny=48, which="SKY", value=0.5)
    correct_ifft_to_lm(a)
    assert np.all(a["SKY"].data == 0.5 * 32 * 48)


def test_make_psf_values_match_oracle(gpu, oracle):
    """make_psf (make_point_spread_function.py:1-79) returns the PSF of
    grid2image_spheroid_ms4(dopsf=True) with pixelincr (-cell_l, cell_m):
    its values equal the oracle's grid2image(do_psf=True) on the same
    weights and flags."""
    from astroviper_amd.make_point_spread_function import make_psf
    cell, ms = cube_dataset(nchan=3, seed=7)
    rng = np.random.default_rng(17)
    ms.WEIGHT.data[...] = rng.uniform(0.5, 1.5, ms.WEIGHT.data.shape)
    ms.FLAG.data[...] = rng.random(ms.FLAG.data.shape) < 0.15
    w = ms.WEIGHT.data.copy()
    flag = ms.FLAG.data.copy()
    psf = make_psf(ms, {"cell_size": (cell, cell), "image_size": (200, 200),
                        "phase_center": (0.0, 0.5), "chan_mode": "cube"},
                   {"sampling": 100, "support": 7})
    p = psf["POINT_SPREAD_FUNCTION"].data
    ref, _, _ = oracle.grid2image(None, ms.UVW.data, w, flag, ms.coords["frequency"].values,
                                  (200, 200), (-cell, cell), chan_mode="cube", do_psf=True)
    assert p.shape == (1,) + ref.shape
    assert np.max(np.abs(p[0] - ref)) <= 1e-9 * np.max(np.abs(ref))
    assert np.all(np.abs(p[0, :, :, 100, 100] - 1 / 0.9999996674 ** 2) < 1e-6)


# ------------------------------------------------- do_imaging_weight mode
@pytest.mark.parametrize("n", [64, 63])
@pytest.mark.parametrize("chan_mode", ["continuum", "cube"])
def test_psf_imaging_weight_matches_oracle(gpu, oracle, n, chan_mode):
    """standard_grid_psf_numpy_wrap with do_imaging_weight (standard_grid.py:
    168-239, 318-430): the weight grid plus its conjugate scatter, through a
    plan of conjugate centres, against the oracle; the 64 grid has conjugate
    footprints that reach past its edge."""
    from test_oracle import imaging_weight_edge_problem
    from astroviper_amd.gcf_prolate_spheroidal import create_prolate_spheroidal_kernel_1D
    from astroviper_amd.standard_grid import standard_grid_psf_numpy_wrap
    uvw, freq, cell, w, n_uv = imaging_weight_edge_problem(n)
    if chan_mode == "cube":  # two channels at f and 1.25 f
        freq = np.array([C, 1.25 * C])
        w = np.concatenate([w, 0.5 * w], axis=2)
    cgk = create_prolate_spheroidal_kernel_1D(100, 7)
    params = {"chan_mode": chan_mode, "image_size_padded": np.array(n_uv), "cell_size": cell,
              "oversampling": 100, "support": 7, "do_psf": True, "do_imaging_weight": True}
    g, sw = standard_grid_psf_numpy_wrap(uvw, w, freq, cgk, params)
    gref, swref = oracle.standard_grid(None, uvw, w, freq, n_uv, cell, chan_mode=chan_mode,
                                       do_psf=True, do_imaging_weight=True, complex_grid=False)
    assert g.dtype == np.float64 and g.shape == gref.shape
    assert rel(g, gref) < 1e-12
    np.testing.assert_allclose(sw, swref, rtol=1e-12)


def test_grid_imaging_weight_with_data_matches_oracle(gpu, oracle):
    """standard_grid_jit with do_imaging_weight on visibilities (complex
    grid): both scatters carry vis * weight."""
    from test_oracle import imaging_weight_edge_problem
    from astroviper_amd.gcf_prolate_spheroidal import create_prolate_spheroidal_kernel_1D
    from astroviper_amd.standard_grid import standard_grid_jit
    uvw, freq, cell, w, n_uv = imaging_weight_edge_problem(64)
    rng = np.random.default_rng(5)
    vis = rng.standard_normal(w.shape) + 1j * rng.standard_normal(w.shape)
    grid = np.zeros((1, w.shape[3]) + n_uv, dtype=np.complex128)
    sw = np.zeros((1, w.shape[3]))
    standard_grid_jit(grid, sw, False, True, vis, uvw, freq, np.zeros(1, dtype=int),
                      np.arange(w.shape[3]), w, create_prolate_spheroidal_kernel_1D(100, 7),
                      np.array(n_uv), cell, 7, 100)
    gref, swref = oracle.standard_grid(vis, uvw, w, freq, n_uv, cell, do_imaging_weight=True)
    assert rel(grid, gref) < 1e-12
    np.testing.assert_allclose(sw, swref, rtol=1e-12)
