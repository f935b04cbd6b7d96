"""The oracle's Hogbom CLEAN restatement (avgrid_oracle.c ora_hogbom_*,
following deconvolvers/hogbom/src/hclean.cpp:51-210) pinned two ways:

* against the reference's own extension, compiled from its sources into
  oracle/_ref by `make -C oracle ref` (skipped where that build is absent):
  bit-identical models, residuals, iteration counts, final peaks and fluxes
  over masks, clean boxes, thresholds, speed-up and start iterations, in
  float32 and float64;
* against what the algorithm fixes by itself: a dirty image that is a
  shifted copy of the PSF (peak 1) is removed in one iteration at gain 1,
  leaving an exactly zero residual; the peak position follows the first
  maximum in row-major order; an iteration cap is honoured.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")


def ref_ext():
    if REF_DIR not in sys.path:
        sys.path.insert(0, REF_DIR)
    return pytest.importorskip("_hogbom_ext")


def gaussian_psf(n, sigma=2.0, dt=np.float64):
    yy, xx = np.mgrid[0:n, 0:n]
    return np.exp(-((xx - n // 2) ** 2 + (yy - n // 2) ** 2) / (2 * sigma ** 2)).astype(dt)


def point_sky(n, rng, k=12, dt=np.float64):
    img = np.zeros((n, n))
    for _ in range(k):
        img[rng.integers(0, n), rng.integers(0, n)] += rng.uniform(-1, 2)
    return img


def dirty_of(sky, psf):
    f = np.fft
    return np.real(f.ifft2(f.fft2(sky) * f.fft2(f.ifftshift(psf))))


CASES = [
    dict(max_iter=300, gain=0.1, threshold=0.05),
    dict(max_iter=200, gain=0.2, threshold=0.0, clean_box=(10, 50, 5, 40)),
    dict(max_iter=400, gain=0.1, threshold=0.02, speedup=50.0),
    dict(max_iter=150, start_iter=30, gain=0.15, threshold=0.0),
    dict(max_iter=250, gain=0.1, threshold=0.01, mask=True),
]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_oracle_hogbom_matches_reference_build(oracle, dt, case):
    H = ref_ext()
    rng = np.random.default_rng(100 + case)
    n = 64
    psf = gaussian_psf(n, 1.7, dt)
    dirty = (dirty_of(point_sky(n, rng), psf) + 0.01 * rng.standard_normal((n, n))).astype(dt)
    kw = dict(CASES[case])
    if kw.pop("mask", False):
        kw["mask"] = (rng.random((n, n)) < 0.7).astype(dt)
    a = oracle.hogbom_clean(dirty, psf, **kw)
    b = H.clean(dirty, psf, **kw)
    assert a["iterations_performed"] == b["iterations_performed"]
    np.testing.assert_array_equal(a["model_image"], b["model_image"])
    np.testing.assert_array_equal(a["residual_image"], b["residual_image"])
    assert a["final_peak"] == b["final_peak"]
    assert a["total_flux_cleaned"] == b["total_flux_cleaned"]
    assert a["converged"] == b["converged"]
    assert a["model_image"].dtype == b["model_image"].dtype == dt
    mk = kw.get("mask")
    assert oracle.hogbom_maximg(dirty, mk) == H.maximg(dirty, mk if mk is not None else
                                                        np.array([], dt))


def test_oracle_hogbom_edge_cases_match_reference_build(oracle):
    H = ref_ext()
    n = 32
    psf = gaussian_psf(n)
    # an all-zero box keeps the previous peak position (0, 0), which lies
    # outside the box and holds a non-zero value
    d = np.zeros((n, n))
    d[0, 0] = 1.0
    for kw in (dict(clean_box=(8, 20, 8, 20), max_iter=5, threshold=0.0),
               dict(clean_box=(8, 20, 8, 20), max_iter=5, threshold=0.1)):
        a, b = oracle.hogbom_clean(d, psf, **kw), H.clean(d, psf, **kw)
        assert a["iterations_performed"] == b["iterations_performed"]
        np.testing.assert_array_equal(a["residual_image"], b["residual_image"])
        np.testing.assert_array_equal(a["model_image"], b["model_image"])
    # NaN pixels are never the peak; ties go to the first in row-major order
    d = np.zeros((n, n))
    d[3, 7] = np.nan
    d[5, 9] = 2.0
    d[5, 2] = -2.0
    a, b = oracle.hogbom_clean(d, psf, max_iter=1), H.clean(d, psf, max_iter=1)
    np.testing.assert_array_equal(a["model_image"], b["model_image"])
    assert a["model_image"][5, 2] == -0.2 and a["model_image"][5, 9] == 0.0


def test_oracle_hogbom_closed_forms(oracle):
    n = 48
    psf = gaussian_psf(n, 2.5)
    # dirty = PSF shifted to (row 20, col 30): one component of exactly 1 at
    # gain 1 removes it; the residual is exactly zero where the shifted PSF
    # overlaps the image
    d = np.roll(np.roll(psf, 20 - n // 2, axis=0), 30 - n // 2, axis=1)
    r = oracle.hogbom_clean(d, psf, max_iter=1, gain=1.0)
    assert r["iterations_performed"] == 1
    assert r["model_image"][20, 30] == 1.0 and np.count_nonzero(r["model_image"]) == 1
    ys, xs = slice(max(0, 20 - n // 2), min(n, 20 + n // 2)), slice(max(0, 30 - n // 2),
                                                                   min(n, 30 + n // 2))
    assert np.all(r["residual_image"][ys, xs] == 0.0)
    # the iteration cap and a threshold above the peak
    r = oracle.hogbom_clean(d, psf, max_iter=7, gain=0.01)
    assert r["iterations_performed"] == 7
    r = oracle.hogbom_clean(d, psf, max_iter=7, gain=0.1, threshold=2.0)
    assert r["iterations_performed"] == 0 and not r["model_image"].any()
    lo, hi = oracle.hogbom_maximg(d)
    assert hi == d.max() and lo == d.min()
