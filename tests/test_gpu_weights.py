"""Imaging weights on the GPU (astroviper_amd.imaging_weighting and
calculate_imaging_weights) against the FP64 oracle: the weight-density grid
and sums of weights (FP64 atomics: equal to the sequential sums to 1e-12),
the Briggs factors, the per-sample imaging weights, and the whole
calculate_imaging_weights flow over two datasets with flags and two or four
polarisations."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

C = 299792458.0


def rel(a, b):
    den = np.nanmax(np.abs(b))
    return float(np.nanmax(np.abs(np.asarray(a) - b)) / (den if den > 0 else 1.0))


def samples(rng, nt=20, nb=120, nchan=4, npol=1, n=96):
    uvw = rng.uniform(-1500, 1500, size=(nt, nb, 3))
    freq = np.linspace(1.0e9, 1.3e9, nchan)
    cell = (-0.5 * C / freq[-1] / 2500.0, 0.5 * C / freq[-1] / 2500.0)
    w = rng.uniform(0.5, 1.5, size=(nt, nb, nchan, npol))
    return uvw, freq, cell, w, (n, n)


@pytest.mark.parametrize("robust", [-2.0, 0.0, 0.5, 2.0])
def test_grid_and_degrid_match_oracle(gpu, oracle, robust):
    from astroviper_amd.imaging_weighting.briggs_weighting import calculate_briggs_params
    from astroviper_amd.imaging_weighting.grid_imaging_weights import (degrid_imaging_weights,
                                                                       grid_imaging_weights)
    rng = np.random.default_rng(int(10 * robust) + 50)
    uvw, freq, cell, w, n_uv = samples(rng)
    w[0, :7] = np.nan   # flagged
    uvw[2, 3, 0] = np.nan
    uvw[1, :4, :2] *= 40  # off the grid
    gp = {"image_size": n_uv, "cell_size": cell}
    g = np.zeros((freq.size, 1) + n_uv)
    s = np.zeros((freq.size, 1))
    grid_imaging_weights(g, s, uvw, w, freq, gp)
    gr = np.zeros_like(g)
    sr = np.zeros_like(s)
    oracle.grid_imaging_weights(gr, sr, uvw, w, freq, n_uv, cell)
    assert rel(g, gr) < 1e-12 and rel(s, sr) < 1e-12
    iwp = {"weighting": "briggs", "robust": robust}
    bf = calculate_briggs_params(g, s, iwp)
    bfr = oracle.calculate_briggs_params(gr, sr, iwp)
    np.testing.assert_allclose(bf, bfr, rtol=1e-12)
    iw = degrid_imaging_weights(g, uvw, w, bf, freq, gp)
    iwr = oracle.degrid_imaging_weights(gr, uvw, w, bfr, freq, n_uv, cell)
    np.testing.assert_array_equal(np.isnan(iw), np.isnan(iwr))
    assert rel(np.nan_to_num(iw), np.nan_to_num(iwr)) < 1e-12
    assert np.all(iw[1, :4] == 0)  # off the grid
    # accumulation into a supplied grid
    grid_imaging_weights(g, s, uvw, w, freq, gp)
    assert rel(g, 2 * gr) < 1e-12 and rel(s, 2 * sr) < 1e-12


def test_nan_briggs_factor_raises(gpu):
    from astroviper_amd.imaging_weighting.grid_imaging_weights import degrid_imaging_weights
    rng = np.random.default_rng(3)
    uvw, freq, cell, w, n_uv = samples(rng, nchan=1)
    g = np.ones((1, 1) + n_uv)
    with pytest.raises(ValueError, match="NaN in briggs_factors"):
        degrid_imaging_weights(g, uvw, w, np.array([[[np.nan]], [[1.0]]]), freq,
                               {"image_size": n_uv, "cell_size": cell})


@pytest.mark.parametrize("npol", [2, 4])
def test_calculate_imaging_weights_flow(gpu, oracle, npol):
    """calculate_imaging_weights.py:28-275 over two datasets: flagged
    samples get NaN, polarisations are averaged (2: 0 and 1; 4: 0 and 3)
    and the result is tiled over the polarisations as WEIGHT_IMAGING."""
    from astroviper_amd.calculate_imaging_weights import calculate_imaging_weights
    from astroviper_amd.msv4 import make_visibility_dataset
    rng = np.random.default_rng(9 + npol)
    dss = {}
    raw = []
    for k in range(2):
        uvw, freq, cell, w, n_uv = samples(rng, nt=10, nb=80, nchan=3, npol=npol)
        flag = rng.random(w.shape) < 0.1
        vis = np.zeros(w.shape, complex)
        dss[f"ms{k}"] = make_visibility_dataset(vis, uvw, w, flag, freq)
        raw.append((uvw, w, flag))
    gp = {"image_size": n_uv, "cell_size": cell}
    ps, group, wgrid = calculate_imaging_weights(dss, gp, {"weighting": "briggs", "robust": 0.5},
                                                 {}, return_weight_density_grid=True)
    assert group["weight_imaging"] == "WEIGHT_IMAGING"
    # the same flow from oracle pieces
    gr = np.zeros((3, 1) + n_uv)
    sr = np.zeros((3, 1))
    singles = []
    for uvw, w, flag in raw:
        ww = w.copy()
        ww[flag] = np.nan
        last = 1 if npol == 2 else 3
        ww = ((ww[..., 0] + ww[..., last]) / 2)[..., None]
        singles.append(ww)
        oracle.grid_imaging_weights(gr, sr, uvw, ww, freq, n_uv, cell)
    assert rel(wgrid, gr) < 1e-12
    bfr = oracle.calculate_briggs_params(gr, sr, {"weighting": "briggs", "robust": 0.5})
    for (uvw, w, flag), ww, name in zip(raw, singles, dss):
        iwr = oracle.degrid_imaging_weights(gr, uvw, ww, bfr, freq, n_uv, cell)
        got = ps[name]["WEIGHT_IMAGING"].data
        assert got.shape == w.shape
        for p in range(npol):
            np.testing.assert_array_equal(np.isnan(got[..., p]), np.isnan(iwr[..., 0]))
            assert rel(np.nan_to_num(got[..., p]), np.nan_to_num(iwr[..., 0])) < 1e-12
    # natural weighting changes nothing
    ps2, g2 = calculate_imaging_weights(dss, gp, {"weighting": "natural"}, {})
    assert g2["weight_imaging"] == "WEIGHT"
