"""Pins for the oracle's restatement of the mosaic gridder's convolution
functions (make_gridding_convolution_function.py, make_pb_symmetric.py):

* the unblocked voltage pattern vanishes at the first zero of J1
  (3.8317059702075125) and the power pattern is its square;
* as the blockage shrinks to zero both obscured formulas tend to 2 J1(r)/r;
* the FFT of a real, centre-symmetric pattern is real and centre-symmetric;
* calc_conv_size on a Gaussian crosses the cut level where the closed form
  says (sigma sqrt(2 ln(1/cut)));
* after resize_and_calc_support every kernel sums to oversampling0 *
  oversampling1 over its support region;
* create_cf_chan_map / create_cf_baseline_map on hand-worked cases;
* the SIN projection in closed form (offsets in Dec or RA alone) and the
  phase gradient at the phase centre (all ones) and one pixel off it.
"""
import numpy as np
import pytest

J1_ZERO = 3.8317059702075125
C = 299792458.0


def grid_params(n=64, cell=4e-4):
    return {"image_size": np.array([n, n]), "image_size_padded": np.array([n, n]),
            "cell_size": np.array([-cell, cell]), "image_center": np.array([n // 2, n // 2])}


def pb_params(dish=10.7, block=0.0, ipower=1):
    return {"list_dish_diameters": np.array([dish]), "list_blockage_diameters": np.array([block]),
            "ipower": ipower}


def test_voltage_pattern_null_and_power(oracle):
    f = 1.4e9
    k = 2 * np.pi * f / C
    a = 10.7 / 2
    cell = J1_ZERO / (k * a)  # pixel (c + 1, c) sits on the first zero
    gp = grid_params(16, cell)
    gp["cell_size"] = np.array([cell, cell])
    v = oracle.gcf_aperture_patterns([f], [0], pb_params(), gp, casa=False)[0, 0, 0]
    c = 8
    assert v[c, c] == 1.0
    assert abs(v[c + 1, c]) < 1e-13 and abs(v[c, c - 1]) < 1e-13
    assert v[c + 1, c + 1] < 0  # past the first zero (r = sqrt(2) x0 < second zero 7.0156)
    p = oracle.gcf_aperture_patterns([f], [0], pb_params(ipower=2), gp, casa=False)[0, 0, 0]
    np.testing.assert_array_equal(p, v * v)


@pytest.mark.parametrize("casa", [False, True])
def test_blockage_limit(oracle, casa):
    gp = grid_params(32, 2e-3)
    ref = oracle.gcf_aperture_patterns([1.2e9], [0], pb_params(), gp, casa=casa)
    tiny = oracle.gcf_aperture_patterns([1.2e9], [0], pb_params(block=1e-4), gp, casa=casa)
    assert np.max(np.abs(tiny - ref)) < 1e-6
    # a real blockage changes the pattern
    big = oracle.gcf_aperture_patterns([1.2e9], [0], pb_params(block=2.0), gp, casa=casa)
    assert np.max(np.abs(big - ref)) > 1e-3


def test_baseline_kernel_is_real_and_symmetric(oracle):
    gp = grid_params(64)
    gcf = {"oversampling": np.array([4, 4]), "list_dish_diameters": np.array([10.7, 6.0]),
           "list_blockage_diameters": np.array([0.75, 0.0]), "ipower": 1}
    bp = oracle.gcf_baseline_patterns([1.4e9], [0], np.array([[0, 0], [0, 1], [1, 1]]), True,
                                      gcf, gp)
    # the (0, 1) pattern is the product of the (0, 0) and (1, 1) square roots
    np.testing.assert_allclose(bp[1] ** 2, bp[0] * bp[2], rtol=1e-12, atol=1e-300)
    f = np.fft
    full = f.fftshift(f.fft2(f.ifftshift(bp[1, 0, 0])))
    assert np.max(np.abs(full.imag)) < 1e-10 * np.max(np.abs(full.real))
    ck = full.real
    c = 32
    sub = ck[c - 20:c + 21, c - 20:c + 21]
    np.testing.assert_allclose(sub, sub[::-1, ::-1], rtol=0, atol=1e-10 * np.abs(ck).max())


def test_calc_conv_size_gaussian(oracle):
    n = 256
    sigma = 5.0
    yy, xx = np.mgrid[0:n, 0:n] - n // 2
    g = np.exp(-(xx ** 2 + yy ** 2) / (2 * sigma ** 2))
    cut = 2.5e-2
    # first pixel at or below the cut: ceil(sigma sqrt(2 ln(1/cut)))
    k = int(np.ceil(sigma * np.sqrt(2 * np.log(1 / cut))))
    ov = np.array([4, 4])
    expect = (int(0.5 + k / 4) + 1) * 2 + 1
    assert oracle.gcf_calc_conv_size(g, np.array([n, n]), cut, ov, np.array([15, 15])) == \
        [expect, expect]
    with pytest.raises(AssertionError):  # beyond max_support
        oracle.gcf_calc_conv_size(g, np.array([n, n]), cut, ov, np.array([expect, expect]))


def test_resize_normalisation(oracle):
    n = 128
    yy, xx = np.mgrid[0:n, 0:n] - n // 2
    ck = np.exp(-(xx ** 2 + yy ** 2) / (2 * 4.0 ** 2))[None, None, None]
    ckc = np.exp(-(xx ** 2 + yy ** 2) / (2 * 6.0 ** 2))[None, None, None]
    gcf = {"resize_conv_size": np.array([64, 64]), "oversampling": np.array([4, 4]),
           "support_cut_level": 2.5e-2, "max_support": np.array([15, 15])}
    rk, rkc, sup = oracle.gcf_resize_and_calc_support(ck, ckc, gcf,
                                                      {"image_size_padded": np.array([n, n])})
    assert rk.shape == (1, 1, 1, 64, 64)
    emb = (sup[0, 0, 0] + 1) * 4
    es = 32 - emb // 2
    for k in (rk, rkc):
        s = k[0, 0, 0, es[0]:es[0] + emb[0], es[1]:es[1] + emb[1]].sum() / 16
        assert abs(s - 1.0) < 1e-12
    # the crop is the centre of the input, scaled
    ratio = rk[0, 0, 0] / ck[0, 0, 0, 32:96, 32:96]
    assert np.ptp(ratio) < 1e-12 * ratio.mean()


def test_chan_map_known_answers(oracle):
    f = np.linspace(1.0e9, 1.4e9, 64)
    cmap, pbf = oracle.gcf_cf_chan_map(f, 0.005)
    # tol = 7 MHz: floor(400 / 7) = 57 bands of 400/57 MHz, centred
    assert len(pbf) == 57
    np.testing.assert_allclose(pbf[0], 1.0e9 + 0.2e9 / 57, rtol=1e-15)
    assert cmap[0] == 0 and cmap[-1] == 56 and np.all(np.diff(cmap) >= 0)
    # more bands than channels: every channel its own
    cmap, pbf = oracle.gcf_cf_chan_map(f[:8], 0.0001)
    np.testing.assert_array_equal(cmap, np.arange(8))
    # a single channel: one band
    cmap, pbf = oracle.gcf_cf_chan_map(np.array([1.2e9]), 0.005)
    np.testing.assert_array_equal(cmap, [0])


def test_baseline_map_known_answer(oracle):
    cbm, pairs = oracle.gcf_cf_baseline_map(np.array([0, 1, 0]),
                                            np.array([[0, 1], [0, 2], [1, 2]]), 2)
    np.testing.assert_array_equal(pairs, [[0, 0], [0, 1], [1, 1]])
    # types (0,1) -> pair 1; (0,0) -> 0; (1,0) matches no i <= j pair and stays 0
    np.testing.assert_array_equal(cbm, [1, 0, 0])


def test_sin_projection_closed_forms(oracle):
    cell = np.array([-1e-4, 1e-4])
    a0, d0 = 1.0, 0.4
    off = oracle.gcf_sin_offset([[a0, d0]], [a0, d0], cell)
    np.testing.assert_array_equal(off, [[0.0, 0.0]])
    dd = 3e-3
    off = oracle.gcf_sin_offset([[a0, d0 + dd]], [a0, d0], cell)
    np.testing.assert_allclose(off, [[0.0, np.sin(dd) / 1e-4]], rtol=1e-12, atol=1e-12)
    da = 2e-3
    off = oracle.gcf_sin_offset([[a0 + da, 0.0]], [a0, 0.0], cell)
    np.testing.assert_allclose(off, [[np.sin(da) / -1e-4, 0.0]], rtol=1e-12, atol=1e-12)


def test_phase_gradient_values(oracle):
    gp = {"image_size_padded": np.array([256, 256]), "cell_size": np.array([-1e-4, 1e-4])}
    gcf = {"phase_direction": np.array([1.0, 0.4]), "oversampling": np.array([10, 10]),
           "resize_conv_size": np.array([40, 40])}
    pg = oracle.gcf_phase_gradient(np.array([[1.0, 0.4]]), gcf, gp)
    assert pg.shape == (1, 40, 40)
    np.testing.assert_array_equal(pg, np.ones_like(pg))
    dd = 1e-3
    pg = oracle.gcf_phase_gradient(np.array([[1.0, 0.4 + dd]]), gcf, gp)
    pix1 = -(np.sin(dd) / 1e-4) * 2 * np.pi / (256 * 10)
    np.testing.assert_allclose(pg[0, 20, 21], np.exp(1j * pix1), rtol=1e-12)
    np.testing.assert_allclose(pg[0, 20, 19], np.exp(-1j * pix1), rtol=1e-12)
    np.testing.assert_allclose(pg[0, 25, 20], 1.0, rtol=1e-12)  # no RA offset
