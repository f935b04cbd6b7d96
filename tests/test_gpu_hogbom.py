"""Hogbom CLEAN on the GPU (astroviper_amd.deconvolvers.hogbom, the
cooperative CLEAN kernel) against the reference's own _hogbom_ext built from
its sources into oracle/_ref, or, where that build is absent, the oracle's
restatement (the two are bit-identical, tests/test_oracle_hogbom.py).

The GPU performs the reference's operations in the same order with the same
roundings, so models, residuals, iteration counts, peaks and fluxes must be
bit-identical; the one exception is the speed-up threshold, 2^x computed by
the device's pow, which is compared through the iteration count."""
import os
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")


class _OracleExt:
    """The oracle's restatement behind the extension's interface."""

    def __init__(self, O):
        self.O = O

    def clean(self, dirty_image, psf, mask=None, **kw):
        kw.pop("progress_callback", None)
        kw.pop("stop_callback", None)
        return self.O.hogbom_clean(dirty_image, psf, mask, **kw)

    def maximg(self, image, mask=None):
        return self.O.hogbom_maximg(image, mask)


@pytest.fixture(scope="module")
def ref(oracle):
    if REF_DIR not in sys.path:
        sys.path.insert(0, REF_DIR)
    try:
        import _hogbom_ext
        return _hogbom_ext
    except ImportError:
        return _OracleExt(oracle)


def gaussian_psf(n, sigma=2.0, dt=np.float64):
    yy, xx = np.mgrid[0:n, 0:n]
    return np.exp(-((xx - n // 2) ** 2 + (yy - n // 2) ** 2) / (2 * sigma ** 2)).astype(dt)


def dirty_of(n, rng, psf, k=20, noise=0.01):
    sky = np.zeros((n, n))
    for _ in range(k):
        sky[rng.integers(0, n), rng.integers(0, n)] += rng.uniform(-1, 2)
    f = np.fft
    d = np.real(f.ifft2(f.fft2(sky) * f.fft2(f.ifftshift(psf))))
    return d + noise * rng.standard_normal((n, n))


def same(a, b, exact_iters=True):
    if exact_iters:
        assert a["iterations_performed"] == b["iterations_performed"]
    np.testing.assert_array_equal(a["model_image"], b["model_image"])
    np.testing.assert_array_equal(a["residual_image"], b["residual_image"])
    assert a["final_peak"] == b["final_peak"]
    assert a["total_flux_cleaned"] == b["total_flux_cleaned"]
    assert a["converged"] == b["converged"]
    assert a["model_image"].dtype == b["model_image"].dtype


CASES = [
    dict(max_iter=300, gain=0.1, threshold=0.05),
    dict(max_iter=200, gain=0.2, threshold=0.0, clean_box=(10, 50, 5, 40)),
    dict(max_iter=400, gain=0.1, threshold=0.02, speedup=50.0),
    dict(max_iter=150, start_iter=30, gain=0.15, threshold=0.0),
    dict(max_iter=250, gain=0.1, threshold=0.01, mask=True),
]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("n", [64, 97])
def test_clean_matches_reference(gpu, ref, dt, case, n):
    from astroviper_amd.deconvolvers import hogbom
    rng = np.random.default_rng(200 + case + n)
    psf = gaussian_psf(n, 1.7, dt)
    dirty = dirty_of(n, rng, psf).astype(dt)
    kw = dict(CASES[case])
    if kw.pop("mask", False):
        kw["mask"] = (rng.random((n, n)) < 0.7).astype(dt)
    a = hogbom.clean(dirty, psf, **kw)
    b = ref.clean(dirty, psf, **kw)
    same(a, b)
    mk = kw.get("mask")
    assert hogbom.maximg(dirty, mk) == ref.maximg(dirty, mk if mk is not None else
                                                  np.array([], dt))


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_clean_large_image(gpu, oracle, dt):
    """1024 x 1024, 300 iterations, against the oracle restatement."""
    from astroviper_amd.deconvolvers import hogbom
    rng = np.random.default_rng(7)
    n = 1024
    psf = gaussian_psf(n, 3.0, dt)
    dirty = dirty_of(n, rng, psf, k=60).astype(dt)
    a = hogbom.clean(dirty, psf, max_iter=300, gain=0.1, threshold=0.0)
    b = oracle.hogbom_clean(dirty, psf, max_iter=300, gain=0.1, threshold=0.0)
    same(a, b)


def test_clean_edge_cases(gpu, ref):
    from astroviper_amd.deconvolvers import hogbom
    n = 32
    psf = gaussian_psf(n)
    d = np.zeros((n, n))
    d[0, 0] = 1.0
    for kw in (dict(clean_box=(8, 20, 8, 20), max_iter=5, threshold=0.0),
               dict(clean_box=(8, 20, 8, 20), max_iter=5, threshold=0.1)):
        same(hogbom.clean(d, psf, **kw), ref.clean(d, psf, **kw))
    d = np.zeros((n, n))
    d[3, 7] = np.nan
    d[5, 9] = 2.0
    d[5, 2] = -2.0
    same(hogbom.clean(d, psf, max_iter=1), ref.clean(d, psf, max_iter=1))
    # a float32 empty mask with float64 images selects the float32 overload,
    # as the reference's deconvolve() calls it (deconvolution.py:399-412)
    rng = np.random.default_rng(3)
    d = dirty_of(n, rng, psf)
    a = hogbom.clean(d, psf, mask=np.array([], dtype=np.float32), max_iter=40)
    assert a["model_image"].dtype == np.float32
    same(a, ref.clean(d, psf, mask=np.array([], dtype=np.float32), max_iter=40))
    same(hogbom.clean_nc(d, psf, 0.2, 0.0, 30), ref.clean(d, psf, gain=0.2, max_iter=30))
    with pytest.raises(RuntimeError, match="PSF dimensions"):
        hogbom.clean(d, psf[:-1])


def test_clean_callbacks(gpu, oracle):
    """The progress callback is called at the start iteration, whenever
    iter % cycle == 1 (cycle = max(1, (max_iter - start_iter) // 10)) and once
    at the end (hclean.cpp:154-160,199-202); a stop callback returning True
    ends the run before that iteration's subtraction."""
    from astroviper_amd.deconvolvers import hogbom
    rng = np.random.default_rng(11)
    n = 64
    psf = gaussian_psf(n)
    d = dirty_of(n, rng, psf)
    calls = []
    a = hogbom.clean(d, psf, max_iter=100, gain=0.1,
                     progress_callback=lambda *x: calls.append(x))
    its = [c[0] for c in calls]
    assert its == [0, 1, 11, 21, 31, 41, 51, 61, 71, 81, 91, 100]
    b = oracle.hogbom_clean(d, psf, max_iter=100, gain=0.1)
    same(a, b)
    # the peak reported at iteration 41 is the peak of the residual after
    # 41 iterations
    r41 = oracle.hogbom_clean(d, psf, max_iter=41, gain=0.1)["residual_image"]
    k = int(np.argmax(np.abs(r41)))
    assert calls[5][1:3] == (k % n, k // n) and calls[5][3] == r41.flat[k]
    # stop at the third report point (iteration 11): 11 iterations done
    asked = []

    def stop():
        asked.append(1)
        return len(asked) == 3  # report points 0, 1, 11

    c = hogbom.clean(d, psf, max_iter=100, gain=0.1, stop_callback=stop)
    assert c["iterations_performed"] == 11 and len(asked) == 3
    same(c, oracle.hogbom_clean(d, psf, max_iter=11, gain=0.1))
