"""The oracle's imaging-weight restatement (avgrid_oracle.c
ora_grid_imaging_weights / ora_degrid_imaging_weights, following
imaging_weighting/grid_imaging_weights.py:105-206,274-443, and
oracle.calculate_briggs_params, briggs_weighting.py:1-31), pinned by what
the definitions fix:

* one sample of weight w on a pixel centre adds w at that pixel and at its
  mirror (N - iu, N - iv), and 2w to the sum of weights;
* the density grid is point-symmetric about the centre and sums to the
  sum of weights when no conjugate falls off the grid;
* robust = +20 gives the natural weights (the Briggs term vanishes), and
  robust = -20 the uniform weights w / (a G) with a = (5e20)^2 sum_w / sum G^2;
* flagged (NaN) weights contribute nothing; off-grid samples get 0.
"""
import numpy as np

C = 299792458.0


def sample_set(rng, nt=4, nb=50, nchan=3, n=64):
    uvw = rng.uniform(-1200, 1200, size=(nt, nb, 3))
    freq = np.linspace(1.0e9, 1.3e9, nchan)
    cell = (-0.5 * C / freq[-1] / 2500.0, 0.5 * C / freq[-1] / 2500.0)
    w = rng.uniform(0.5, 1.5, size=(nt, nb, nchan, 1))
    return uvw, freq, cell, w, (n, n)


def test_single_sample_and_mirror(oracle):
    n = 64
    cell = (-1e-4, 1e-4)
    f = 1.0e9
    su = -(f * cell[0] * n) / C
    sv = -(f * cell[1] * n) / C
    uvw = np.zeros((1, 1, 3))
    uvw[0, 0, 0] = 5.0 / su   # u position 37 (n//2 + 5)
    uvw[0, 0, 1] = -7.0 / sv  # v position 25
    g = np.zeros((1, 1, n, n))
    s = np.zeros((1, 1))
    oracle.grid_imaging_weights(g, s, uvw, np.full((1, 1, 1, 1), 1.5), np.array([f]), (n, n), cell)
    assert g[0, 0, 37, 25] == 1.5 and g[0, 0, 27, 39] == 1.5 and g.sum() == 3.0
    assert s[0, 0] == 3.0


def test_density_grid_symmetry_and_sum(oracle):
    rng = np.random.default_rng(1)
    uvw, freq, cell, w, n_uv = sample_set(rng)
    w[0, :5] = np.nan
    g = np.zeros((freq.size, 1) + n_uv)
    s = np.zeros((freq.size, 1))
    oracle.grid_imaging_weights(g, s, uvw, w, freq, n_uv, cell)
    n = n_uv[0]
    # cell (i, j) mirrors to (n - i, n - j); row/column 0 has no mirror
    core = g[:, 0, 1:, 1:]
    np.testing.assert_allclose(core, core[:, ::-1, ::-1], rtol=0, atol=1e-12)
    # every conjugate landed inside (no sample on row or column 0), so the
    # grid holds exactly the 2 w of every gridded sample
    assert not g[:, 0, 0, :].any() and not g[:, 0, :, 0].any()
    np.testing.assert_allclose(g.sum(axis=(1, 2, 3)), s[:, 0], rtol=1e-12)
    assert np.all(np.isfinite(g))


def test_briggs_limits(oracle):
    rng = np.random.default_rng(2)
    uvw, freq, cell, w, n_uv = sample_set(rng, nt=6, nb=80)
    g = np.zeros((freq.size, 1) + n_uv)
    s = np.zeros((freq.size, 1))
    oracle.grid_imaging_weights(g, s, uvw, w, freq, n_uv, cell)
    on = np.zeros(w.shape, bool)
    nat = oracle.calculate_briggs_params(g, s, {"weighting": "briggs", "robust": 20.0})
    iw = oracle.degrid_imaging_weights(g, uvw, w, nat, freq, n_uv, cell)
    on = iw != 0
    assert on.mean() > 0.5
    np.testing.assert_allclose(iw[on], w[on], rtol=1e-12)
    uni = oracle.calculate_briggs_params(g, s, {"weighting": "briggs", "robust": -20.0})
    iwu = oracle.degrid_imaging_weights(g, uvw, w, uni, freq, n_uv, cell)
    a = (5.0 * 1e20) ** 2 / (np.sum(g ** 2, axis=(2, 3)) / s)
    # the uniform limit: w / (a G), with G the density at the sample's pixel
    nt, nb, nchan, _ = w.shape
    for ch in range(nchan):
        su = -(freq[ch] * cell[0] * n_uv[0]) / C
        sv = -(freq[ch] * cell[1] * n_uv[1]) / C
        iu = (uvw[..., 0] * su + n_uv[0] // 2 + 0.5).astype(int)
        iv = (uvw[..., 1] * sv + n_uv[1] // 2 + 0.5).astype(int)
        ok = (iu >= 0) & (iu < n_uv[0]) & (iv >= 0) & (iv < n_uv[1])
        G = g[ch, 0, np.clip(iu, 0, n_uv[0] - 1), np.clip(iv, 0, n_uv[1] - 1)]
        expect = np.where(ok, w[:, :, ch, 0] / (a[ch, 0] * G + 1.0), 0.0)
        np.testing.assert_allclose(iwu[:, :, ch, 0], expect, rtol=1e-12)
    # briggs_abs and natural factor layouts
    ab = oracle.calculate_briggs_params(g, s, {"weighting": "briggs_abs", "robust": 0.5,
                                               "briggs_abs_noise": 2.0})
    assert np.all(ab[0] == 0.25) and np.all(ab[1] == 8.0)
    nn = oracle.calculate_briggs_params(g, s, {"weighting": "natural"})
    assert nn.shape == (2, 1, 1, freq.size, 1) and np.all(nn[0] == 1) and np.all(nn[1] == 0)
