"""Pins the oracle's mosaic gridder (mosaic_grid_jit, mosaic_grid.py:306-486)
before the GPU is compared with it (CPU only).

* With the standard kernel written out as a 2-D oversampled convolution
  function (cf[x, y] = cgk[|x - c|] * cgk[|y - c|]) and a unit phase
  gradient, its grid must be the standard gridder's on samples that both
  bounds tests keep, its PSF sum of weights the standard one, and its data
  sum of weights weight * norm^2 per sample.
* A constant phase factor e^{i phi} on the convolution function multiplies
  the grid by e^{i phi} and the PSF sum of weights by cos(phi).
* On a tiny problem with a random complex convolution function, per-axis
  oversampling and supports that differ by baseline group, channel and
  polarisation, it matches a cell-centric direct sum.
"""
import numpy as np
import pytest

C = 299792458.0


def inner_problem(rng, nt=3, nb=30, nchan=2, npol=2, n=64, margin=9):
    """Samples at f = c with 1/n cells (pixel positions = -uvw), all at
    least `margin` pixels inside the grid."""
    uvw = np.zeros((nt, nb, 3))
    uvw[..., :2] = rng.uniform(-(n / 2 - margin), n / 2 - margin, size=(nt, nb, 2))
    freq = np.full(nchan, C) * (1 + 0.0 * np.arange(nchan))
    vis = rng.standard_normal((nt, nb, nchan, npol)) + 1j * rng.standard_normal((nt, nb, nchan, npol))
    w = rng.uniform(0.5, 1.5, size=(nt, nb, nchan, npol))
    return uvw, freq, (1.0 / n, 1.0 / n), vis, w, (n, n)


def ps_cf(oracle, ov=100, support=7):
    cgk = oracle.create_prolate_spheroidal_kernel_1D(ov, support)
    c = ov * (support // 2) + ov // 2
    x = np.abs(np.arange(2 * c + 1) - c)
    k1 = cgk[x]
    return np.outer(k1, k1)[None, None, None].astype(np.complex128)


def call_mosaic(oracle, psf, vis, uvw, w, freq, n_uv, cell, cf, ws, ov, pg=None, cube=False):
    nt, nb, nchan, npol = w.shape
    nic = nchan if cube else 1
    grid = np.zeros((nic, npol) + n_uv, dtype=np.complex128)
    sw = np.zeros((nic, npol))
    if pg is None:
        pg = np.ones((1,) + cf.shape[-2:], dtype=np.complex128)
    chan_map = np.arange(nchan) if cube else np.zeros(nchan, int)
    oracle.mosaic_grid(grid, sw, psf, vis, uvw, freq, chan_map, np.arange(npol),
                       np.zeros(nb, int), np.zeros(nchan, int), np.zeros(npol, int), w, cf,
                       np.array(n_uv), cell, ws, np.array([ov, ov]), pg)
    return grid, sw


def test_separable_kernel_reduces_to_the_standard_gridder(oracle):
    rng = np.random.default_rng(31)
    uvw, freq, cell, vis, w, n_uv = inner_problem(rng)
    cf = ps_cf(oracle)
    ws = np.full((1, 1, 1, 2), 7)
    for psf in (True, False):
        g, sw = call_mosaic(oracle, psf, vis, uvw, w, freq, n_uv, cell, cf, ws, 100)
        gs, sws = oracle.standard_grid(None if psf else vis, uvw, w, freq, n_uv, cell,
                                       do_psf=psf)
        assert np.max(np.abs(g - gs)) <= 1e-13 * np.max(np.abs(gs))
        if psf:
            np.testing.assert_allclose(sw, sws, rtol=1e-13)
    # the data sum of weights is sum w * norm^2, norm the standard per-sample sum
    expect = np.zeros_like(sw)
    for t in range(uvw.shape[0]):
        for b in range(uvw.shape[1]):
            for c in range(freq.size):
                _, s1 = oracle.standard_grid(None, uvw[t:t + 1, b:b + 1], np.ones((1, 1, 1, 2)),
                                             freq[c:c + 1], n_uv, cell, do_psf=True)
                expect[0] += w[t, b, c] * s1[0, 0] ** 2
    np.testing.assert_allclose(sw, expect, rtol=1e-12)


def test_constant_phase_rotates_grid(oracle):
    rng = np.random.default_rng(32)
    uvw, freq, cell, vis, w, n_uv = inner_problem(rng)
    cf = ps_cf(oracle)
    ws = np.full((1, 1, 1, 2), 7)
    phi = 0.7
    pg = np.full((2,) + cf.shape[-2:], np.exp(1j * phi))  # field 0 is used
    g0, s0 = call_mosaic(oracle, True, vis, uvw, w, freq, n_uv, cell, cf, ws, 100)
    g1, s1 = call_mosaic(oracle, True, vis, uvw, w, freq, n_uv, cell, cf, ws, 100, pg=pg)
    assert np.max(np.abs(g1 - np.exp(1j * phi) * g0)) <= 1e-13 * np.max(np.abs(g0))
    np.testing.assert_allclose(s1, np.cos(phi) * s0, rtol=1e-13)


def test_matches_cell_centric_sum(oracle):
    rng = np.random.default_rng(33)
    n, nt, nb, nchan, npol = 40, 2, 6, 2, 2
    ov = (4, 6)
    ws = np.array([[[[5, 7], [3, 5]], [[7, 7], [5, 3]]],
                   [[[3, 3], [5, 5]], [[7, 5], [3, 7]]]])  # [cfb][cfc][cfp][2]
    smax = int(ws.max())
    cu_n, cv_n = ov[0] * (smax + 1) + 1, ov[1] * (smax + 1) + 1
    cf = (rng.standard_normal((2, 2, 2, cu_n, cv_n))
          + 1j * rng.standard_normal((2, 2, 2, cu_n, cv_n)))
    pg = np.exp(1j * rng.uniform(0, 6.3, size=(1, cu_n, cv_n)))
    uvw = np.zeros((nt, nb, 3))
    uvw[..., :2] = rng.uniform(-(n / 2 - 9), n / 2 - 9, size=(nt, nb, 2))
    freq = np.array([C, 1.1 * C])
    cell = (1.0 / n, 1.0 / n)
    vis = rng.standard_normal((nt, nb, nchan, npol)) + 1j * rng.standard_normal((nt, nb, nchan, npol))
    w = rng.uniform(0.5, 1.5, size=(nt, nb, nchan, npol))
    cfbm, cfcm, cfpm = np.array([0, 1, 1, 0, 1, 0]), np.array([1, 0]), np.array([0, 1])
    for psf in (False, True):
        grid = np.zeros((nchan, npol, n, n), dtype=np.complex128)
        sw = np.zeros((nchan, npol))
        oracle.mosaic_grid(grid, sw, psf, vis, uvw, freq, np.arange(nchan), np.arange(npol),
                           cfbm, cfcm, cfpm, w, cf, np.array([n, n]), cell, ws, np.array(ov), pg)
        # cell-centric: each cell sums the cf value at its offset from every sample
        cfpg = cf * pg[0]
        gb = np.zeros_like(grid)
        swb = np.zeros_like(sw)
        uu, vv = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        for t in range(nt):
            for b in range(nb):
                for c in range(nchan):
                    s = -(freq[c] * cell[0] * n) / C
                    up, vp = uvw[t, b, 0] * s + n // 2, uvw[t, b, 1] * s + n // 2
                    cu, cv = int(up + 0.5), int(vp + 0.5)
                    if not (cu + smax < n and cv + smax < n and cu - smax >= 0 and cv - smax >= 0):
                        continue
                    ou = int(np.floor((cu - up) * ov[0] + 0.5)) + cu_n // 2
                    ovv = int(np.floor((cv - vp) * ov[1] + 0.5)) + cv_n // 2
                    for p in range(npol):
                        su, sv = ws[cfbm[b], cfcm[c], cfpm[p]]
                        du, dv = uu - cu, vv - cv
                        inside = ((du >= -(su // 2)) & (du < su - su // 2) & (dv >= -(sv // 2))
                                  & (dv < sv - sv // 2))
                        kern = np.where(inside, cfpg[cfbm[b], cfcm[c], cfpm[p]][
                            np.clip(ov[0] * du + ou, 0, cu_n - 1),
                            np.clip(ov[1] * dv + ovv, 0, cv_n - 1)], 0)
                        d = w[t, b, c, p] if psf else vis[t, b, c, p] * w[t, b, c, p]
                        gb[c, p] += kern * d
                        norm = kern.sum()
                        swb[c, p] += w[t, b, c, p] * (norm.real if psf else (norm * norm).real)
        assert np.max(np.abs(grid - gb)) <= 1e-12 * np.max(np.abs(gb))
        np.testing.assert_allclose(sw, swb, rtol=1e-12)
