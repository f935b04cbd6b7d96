"""Device-resident major cycles (astroviper_amd.major_cycle) against the
FP64 oracle step by step -- gridding of VISIBILITY and RESIDUAL
(imager.py:54-79), the PSF, the model prediction (imager.py:88-105: FFT of
the model without taper, then degridding), the residual visibilities
(imager.py:82-85) and the per-plane Hogbom CLEAN (deconvolution.py:224-287,
float32 as the reference's hogbom_clean selects) -- and over two whole
major cycles against the same loop built from oracle pieces."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

C = 299792458.0


def rel(a, b):
    den = np.max(np.abs(b))
    return float(np.max(np.abs(np.asarray(a) - b)) / (den if den > 0 else 1.0))


def dataset(seed=5, n_time=40, nchan=2, npol=2, n=128):
    """27 antennas, 5 point sources inside a 128 x 128 field of 5" cells,
    10% flags, weights U(0.5, 1.5), two NaN-uvw rows."""
    from astroviper_amd import synth
    from astroviper_amd.msv4 import make_visibility_dataset
    arr = synth.ArrayConfig(n_time=n_time, n_chan=nchan, n_pol=npol, seed=seed)
    sky = synth.SkyConfig(n_src=5, flux=(0.5, 2.0), max_offset_deg=0.05, noise=0.005)
    obs = synth.make_observation(arr, sky)
    vis, w, flag = synth.host_visibilities(obs, sky, flag_frac=0.1, weight_range=(0.5, 1.5))
    uvw = obs.uvw.copy()
    uvw[3, 10, 0] = np.nan
    uvw[7, 2, 1] = np.nan
    cell = 5.0 * synth.ARCSEC
    ms = make_visibility_dataset(vis, uvw, w, flag, obs.freq)
    return ms, (n, n), (-cell, cell)


def oracle_dirty(oracle, ms, vis, n_uv, cell, psf=False):
    img, g, sw = oracle.grid2image(vis, ms.UVW.data, ms.WEIGHT.data, ms.FLAG.data,
                                   ms.coords["frequency"].values, n_uv, cell, chan_mode="cube",
                                   do_psf=psf)
    return img


def oracle_predict(oracle, ms, model, cell):
    grid = oracle.fft_lm_to_uv(model, axes=[2, 3])
    return oracle.degrid(ms.UVW.data, ms.FLAG.data, ms.coords["frequency"].values, grid, cell)


def test_resident_steps_match_oracle(gpu, oracle):
    from astroviper_amd.major_cycle import ResidentImager
    ms, n_uv, cell = dataset()
    im = ResidentImager(ms, n_uv, cell, chan_mode="cube")
    vis = ms.VISIBILITY.data
    dirty = im.grid_visibilities("VISIBILITY").cpu().numpy()
    dref = oracle_dirty(oracle, ms, vis, n_uv, cell)
    assert rel(dirty, dref) < 1e-9
    psf = im.make_psf().cpu().numpy()
    pref = oracle_dirty(oracle, ms, vis, n_uv, cell, psf=True)
    assert rel(psf, pref) < 1e-9
    # a sparse model predicts the oracle's model visibilities; flagged and
    # NaN-uvw samples get 0
    rng = np.random.default_rng(1)
    model = np.zeros(dirty.shape)
    for _ in range(10):
        model[rng.integers(0, 2), rng.integers(0, 2), rng.integers(40, 88),
              rng.integers(40, 88)] = rng.uniform(0.1, 1.0)
    im.predict_model_visibilities(model)
    mv = im.model_vis.cpu().numpy().reshape(vis.shape)
    mref = oracle_predict(oracle, ms, model, cell)
    assert rel(mv, mref) < 1e-12
    assert np.all(mv[ms.FLAG.data] == 0) and np.all(mv[3, 10] == 0)
    im.compute_residual_visibilities()
    rv = im.residual_vis.cpu().numpy().reshape(vis.shape)
    np.testing.assert_array_equal(rv, vis - mv)
    # the residual visibilities grid to the oracle's residual image
    r2 = im.grid_visibilities("RESIDUAL").cpu().numpy()
    assert rel(r2, oracle_dirty(oracle, ms, vis - mref, n_uv, cell)) < 1e-9
    # CLEAN of every plane equals the oracle's float32 CLEAN of the same
    # dirty plane with the channel's PSF
    comps, resid, iters = im.deconvolve(im.grid_visibilities("VISIBILITY"), im.make_psf(), 60,
                                        0.1, 0.0)
    d32 = dirty.astype(np.float32)
    p32 = psf.astype(np.float32)
    for c in range(dirty.shape[0]):
        for q in range(dirty.shape[1]):
            r = oracle.hogbom_clean(d32[c, q], p32[c, 0], np.array([], np.float32),
                                    max_iter=60, gain=0.1)
            np.testing.assert_array_equal(comps[c, q].cpu().numpy(), r["model_image"])
            np.testing.assert_array_equal(resid[c, q].cpu().numpy(), r["residual_image"])
            assert iters[c, q] == r["iterations_performed"]
    im.write_back(ms)
    assert ms["VISIBILITY_MODEL"].data.dtype == np.complex128
    np.testing.assert_array_equal(ms["RESIDUAL"].data, rv)
    im.close()


def test_major_cycles_match_oracle_loop(gpu, oracle):
    from astroviper_amd.major_cycle import run_major_cycles
    ms, n_uv, cell = dataset(seed=6)
    vis = ms.VISIBILITY.data
    model, resid, psf, log = run_major_cycles(ms, n_uv, cell, nmajor=2, cycle_niter=80, gain=0.1)
    # the same loop from oracle pieces
    psf_ref = oracle_dirty(oracle, ms, vis, n_uv, cell, psf=True).astype(np.float32)
    m_ref = np.zeros(model.shape)
    r_vis = vis.copy()
    for _ in range(2):
        dirty = oracle_dirty(oracle, ms, r_vis, n_uv, cell).astype(np.float32)
        r_img = np.zeros(dirty.shape)
        for c in range(dirty.shape[0]):
            for q in range(dirty.shape[1]):
                r = oracle.hogbom_clean(dirty[c, q], psf_ref[c, 0], np.array([], np.float32),
                                        max_iter=80, gain=0.1)
                m_ref[c, q] += r["model_image"]
                r_img[c, q] = r["residual_image"]
        r_vis = vis - oracle_predict(oracle, ms, m_ref, cell)
    print("major cycles:", log)
    assert len(log) == 2 and log[1]["peak_before"] < log[0]["peak_before"]
    assert rel(model.cpu().numpy(), m_ref) < 1e-5
    assert rel(resid.cpu().numpy(), r_img) < 1e-5


@pytest.mark.parametrize("shape", [(3, 64, 128), (2, 2, 36, 20), (2, 60, 66), (1, 62, 64)])
@pytest.mark.parametrize("cdt", ["complex64", "complex128"])
def test_model_to_uv_matches_shifted_fft(gpu, shape, cdt):
    """model_to_uv (the sign-flip form S * fft2(S * x) when both axes are
    multiples of 4, the two shifts otherwise) against NumPy's
    fftshift(fft2(ifftshift(x))), the reference's fft_lm_to_uv (fft.py:5-18),
    in FP64 on the CPU."""
    from astroviper_amd.major_cycle import model_to_uv
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape)
    ax = (-2, -1)
    ref = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(x, axes=ax), axes=ax), axes=ax)
    got = model_to_uv(torch.from_numpy(x).to(gpu), getattr(torch, cdt)).cpu().numpy()
    assert got.dtype == np.dtype(cdt)
    assert rel(got, ref) < (1e-6 if cdt == "complex64" else 1e-13)
    # a complex image takes the complex FFT
    z = x + 1j * rng.standard_normal(shape)
    ref = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(z, axes=ax), axes=ax), axes=ax)
    got = model_to_uv(torch.from_numpy(z).to(gpu), getattr(torch, cdt)).cpu().numpy()
    assert rel(got, ref) < (1e-6 if cdt == "complex64" else 1e-13)
