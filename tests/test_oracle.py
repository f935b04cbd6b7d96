"""Pins the FP64 CPU oracle (oracle/) before any GPU result is compared with it.

Each check is chosen so that a plausible restatement mistake -- a wrong
coefficient, a dropped term, a transposed axis, a wrong rounding rule or sign
-- fails: first-principles spheroidal functions, closed forms, invariants,
brute force on tiny inputs and the reference test-suite's own known answers
(tests/unit/core/imaging/...).  CPU only.
"""
import json
import math
import os

import numpy as np
import pytest

C = 299792458.0
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")


def spheroidal_alpha1(eta, c=3 * np.pi, n=96):
    """psi_{alpha=1}(c, eta), psi(0) = 1, from the Nystrom solution of
    lambda psi(x) = int_-1^1 (1-t^2) cos(c x t) psi(t) dt (Gauss-Legendre)."""
    t, w = np.polynomial.legendre.leggauss(n)
    d = w * (1 - t ** 2)
    S = np.sqrt(d)[:, None] * np.cos(c * t[:, None] * t[None, :]) * np.sqrt(d)[None, :]
    lam, V = np.linalg.eigh(S)
    v = V[:, -1] / np.sqrt(d)
    eta = np.atleast_1d(eta)
    vals = (np.cos(c * eta[:, None] * t[None, :]) * (d * v)[None, :]).sum(1)
    return vals / (d * v).sum()


# ------------------------------------------------------------ kernel table
def test_grdsf_matches_first_principles_spheroidal(oracle):
    eta = np.linspace(0.0, 1.0, 401)
    g, k = oracle.prolate_spheroidal_function(eta)
    ref = spheroidal_alpha1(eta)
    # Schwab's rational fit is accurate to ~1e-7 of the peak
    assert np.max(np.abs(g - ref)) < 1e-6
    np.testing.assert_allclose(k, (1 - eta ** 2) * g, rtol=0, atol=1e-15)


def test_grdsf_closed_forms_and_continuity(oracle):
    g, k = oracle.prolate_spheroidal_function(np.array([1.0, 0.75, 0.75 - 1e-12, 1.5, -0.3, 0.3]))
    assert g[0] == 4.028559e-3                  # d = 0 in the second segment: p[1,0]/q[1,0]
    assert abs(g[1] - g[2]) < 2e-7              # the two segments meet at 0.75
    assert g[3] == 0.0 and k[3] == 0.0          # |u| > 1
    assert g[4] == g[5]                          # even function
    assert k[0] == 0.0                           # (1 - u^2) vanishes at the edge
    # value at 0 from the first segment, d = -0.5625 (gcf_prolate_spheroidal.py:116-127)
    g0, _ = oracle.prolate_spheroidal_function(np.array([0.0]))
    assert abs(g0[0] - 0.9999996674) < 1e-9


def test_kernel_table(oracle):
    cgk = oracle.create_prolate_spheroidal_kernel_1D(100, 7)
    assert cgk.shape == (400,)
    assert np.all(cgk[300:] == 0.0) and np.all(cgk[:300] > 0.0)
    assert np.all(np.diff(cgk[:300]) < 0)       # monotone decreasing
    u = np.arange(300) / 300.0
    ref = (1 - u ** 2) * spheroidal_alpha1(u)
    assert np.max(np.abs(cgk[:300] - ref)) < 1e-6
    with open(GOLDEN) as f:
        gold = json.load(f)["kernel_table_support7_ov100"]
    for i, v in gold.items():
        assert abs(cgk[int(i)] - v) < 5e-8, i
    # support 5 and 9 tables (config 4)
    for s in (5, 9):
        t = oracle.create_prolate_spheroidal_kernel_1D(100, s)
        assert t.shape == (100 * (s // 2 + 1),)
        assert np.all(t[100 * (s // 2):] == 0.0)


def test_taper_is_outer_product_of_grdsf(oracle):
    _, corr = oracle.create_prolate_spheroidal_kernel(100, 7, np.array([64, 48]))
    gu, _ = oracle.prolate_spheroidal_function(np.abs(2.0 * oracle.coordinates(64)))
    gv, _ = oracle.prolate_spheroidal_function(np.abs(2.0 * oracle.coordinates(48)))
    np.testing.assert_array_equal(corr, np.outer(gu, gv))
    assert corr.min() > 0                        # never zero: grdsf(1) = 4.03e-3
    assert abs(corr[32, 24] - 0.9999996674 ** 2) < 1e-9


def test_four_d_kernel_structure(oracle):
    k4, _ = oracle.create_prolate_spheroidal_kernel(100, 7, np.array([8, 8]))
    assert k4.shape == (101, 101, 7, 7)
    assert np.all(k4[100] == 0) and np.all(k4[:, 100] == 0)  # loops run to ov only
    # row 50 is the zero-offset row: its centre tap is grdsf(0)
    assert abs(k4[50, 50, 3, 3] - 0.9999996674 ** 2) < 1e-9


# --------------------------------------------------------------- geometry
def test_degrid_geometry_known_answers(oracle):
    """test_standard_degrid.py:51-121: pos, loc = round-half-even, off."""
    with open(GOLDEN) as f:
        gold = json.load(f)["sgrid_basic"]
    pos, loc, off = oracle.degrid_geometry(gold["uvw"], gold["freq"], gold["scale"],
                                           gold["offset"], gold["sampling"])
    ex = gold["scale"][0] * gold["uvw"][0] * gold["freq"] / C + gold["offset"][0]
    ey = gold["scale"][1] * gold["uvw"][1] * gold["freq"] / C + gold["offset"][1]
    assert pos[0] == ex and pos[1] == ey
    assert list(loc) == [int(round(ex)), int(round(ey))]
    assert list(off) == [int(round((round(ex) - ex) * gold["sampling"])),
                         int(round((round(ey) - ey) * gold["sampling"]))]


def test_rounding_rules_at_exact_ties(oracle):
    # f = c, dl = 1/N, N power of two => scale = -1 exactly, u_pos = N/2 - uvw_u
    n = 64
    for x in (0.5, 1.5, 2.5, -3.5, 7.5):
        g = oracle.grid_geometry(x, -x, C, (1.0 / n, 1.0 / n), (n, n))
        # grid: int(u_pos + 0.5) rounds half up
        assert g[0] == int(n / 2 - x + 0.5) == math.floor(n / 2 - x + 0.5)
        assert g[1] == int(n / 2 + x + 0.5)
        assert g[2] == math.floor((g[0] - (n / 2 - x)) * 100 + 0.5)
        pos, loc, off = oracle.degrid_geometry([x, -x, 0], C, [-1.0, -1.0], [n / 2, n / 2], 100)
        assert pos[0] == n / 2 - x
        assert loc[0] == round(n / 2 - x)        # Python round: half to even
        assert loc[1] == round(n / 2 + x)
    # the two rules differ exactly at half-integers with an even floor
    g = oracle.grid_geometry(0.5, 0.0, C, (1.0 / n, 1.0 / n), (n, n))
    _, loc, _ = oracle.degrid_geometry([0.5, 0, 0], C, [-1.0, -1.0], [n / 2, n / 2], 100)
    assert g[0] == 32 and loc[0] == 32          # 31.5: up -> 32, even -> 32
    g = oracle.grid_geometry(-0.5, 0.0, C, (1.0 / n, 1.0 / n), (n, n))
    _, loc, _ = oracle.degrid_geometry([-0.5, 0, 0], C, [-1.0, -1.0], [n / 2, n / 2], 100)
    assert g[0] == 33 and loc[0] == 32          # 32.5: up -> 33, even -> 32


# ------------------------------------------------------------------ gridding
def _random_problem(rng, nt=3, nb=40, nchan=3, npol=2, n=64, cell=None, frac_flag=0.0):
    uvw = rng.uniform(-1500.0, 1500.0, size=(nt, nb, 3))
    freq = np.linspace(1.0e9, 1.3e9, nchan)
    if cell is None:
        cell = (-0.5 * C / freq[-1] / 2500.0, 0.5 * C / freq[-1] / 2500.0)
    vis = rng.standard_normal((nt, nb, nchan, npol)) + 1j * rng.standard_normal((nt, nb, nchan, npol))
    w = rng.uniform(0.5, 1.5, size=(nt, nb, nchan, npol))
    flag = rng.random((nt, nb, nchan, npol)) < frac_flag
    return uvw, freq, cell, vis, w, flag, (n, n)


def brute_force_grid(oracle, vis, uvw, w, freq, cell, n_uv, support=7, ov=100, cube=False,
                     psf=False):
    """Cell-centric direct sum: every cell sums the kernel weight of every
    sample whose support covers it (no scatter loop)."""
    nt, nb, nchan, npol = w.shape
    nic = nchan if cube else 1
    grid = np.zeros((nic, npol, n_uv[0], n_uv[1]), dtype=np.complex128)
    sw = np.zeros((nic, npol))
    cgk = oracle.create_prolate_spheroidal_kernel_1D(ov, support)
    sc = support // 2
    uu, vv = np.meshgrid(np.arange(n_uv[0]), np.arange(n_uv[1]), indexing="ij")
    for t in range(nt):
        for b in range(nb):
            for c in range(nchan):
                geo = oracle.grid_geometry(uvw[t, b, 0], uvw[t, b, 1], freq[c], cell, n_uv,
                                           support, ov)
                if geo is None:
                    continue
                iu, iv, ou, ovv = geo
                du, dv = uu - iu, vv - iv
                inside = (du >= -sc) & (du < support - sc) & (dv >= -sc) & (dv < support - sc)
                ku = cgk[np.abs(ov * np.clip(du, -sc, sc) + ou)]
                kv = cgk[np.abs(ov * np.clip(dv, -sc, sc) + ovv)]
                kern = np.where(inside, ku * kv, 0.0)
                for p in range(npol):
                    d = w[t, b, c, p] if psf else vis[t, b, c, p] * w[t, b, c, p]
                    if d == 0 or np.isnan(d):
                        continue
                    ic = c if cube else 0
                    grid[ic, p] += kern * d
                    sw[ic, p] += w[t, b, c, p] * kern.sum()
    return grid, sw


@pytest.mark.parametrize("cube", [False, True])
def test_grid_matches_brute_force(oracle, cube):
    rng = np.random.default_rng(11)
    uvw, freq, cell, vis, w, flag, n_uv = _random_problem(rng)
    g, sw = oracle.standard_grid(vis, uvw, w, freq, n_uv, cell,
                                 chan_mode="cube" if cube else "continuum")
    gb, swb = brute_force_grid(oracle, vis, uvw, w, freq, cell, n_uv, cube=cube)
    assert np.max(np.abs(g - gb)) <= 1e-12 * np.max(np.abs(gb))
    np.testing.assert_allclose(sw, swb, rtol=1e-12)


def brute_force_imaging_weight_psf(oracle, uvw, w, freq, cell, n_uv, support=7, ov=100):
    """Cell-centric sum of do_imaging_weight's PSF (standard_grid.py:318-430)
    written from its definition: every sample adds its kernel around its grid
    centre and, with the same kernel weights, around its conjugate centre
    int(-u + N/2 + 0.5) (cells outside the grid dropped); for two or more
    polarisations the gridded value is (w0 + w1) / 2 for every polarisation,
    and the sum of weights counts both scatters."""
    nt, nb, nchan, npol = w.shape
    grid = np.zeros((1, npol, n_uv[0], n_uv[1]))
    sw = np.zeros((1, npol))
    cgk = oracle.create_prolate_spheroidal_kernel_1D(ov, support)
    sc = support // 2
    uu, vv = np.meshgrid(np.arange(n_uv[0]), np.arange(n_uv[1]), indexing="ij")

    def footprint(cu, cv, ou, ovv):
        du, dv = uu - cu, vv - cv
        inside = (du >= -sc) & (du < support - sc) & (dv >= -sc) & (dv < support - sc)
        ku = cgk[np.abs(ov * np.clip(du, -sc, sc) + ou)]
        kv = cgk[np.abs(ov * np.clip(dv, -sc, sc) + ovv)]
        return np.where(inside, ku * kv, 0.0)

    for t in range(nt):
        for b in range(nb):
            for c in range(nchan):
                geo = oracle.grid_geometry(uvw[t, b, 0], uvw[t, b, 1], freq[c], cell, n_uv,
                                           support, ov)
                if geo is None:
                    continue
                iu, iv, ou, ovv = geo
                u = uvw[t, b, 0] * (-(freq[c] * cell[0] * n_uv[0]) / C)
                v = uvw[t, b, 1] * (-(freq[c] * cell[1] * n_uv[1]) / C)
                cu, cv = int(-u + n_uv[0] // 2 + 0.5), int(-v + n_uv[1] // 2 + 0.5)
                k0 = footprint(iu, iv, ou, ovv)
                k1 = footprint(cu, cv, ou, ovv)
                norm = k0.sum()
                for p in range(npol):
                    d = (w[t, b, c, 0] + w[t, b, c, 1]) / 2.0 if npol >= 2 else w[t, b, c, p]
                    if d == 0 or np.isnan(d):
                        continue
                    grid[0, p] += (k0 + k1) * d
                    sw[0, p] += 2 * d * norm
    return grid, sw


def imaging_weight_edge_problem(n):
    """Samples at f = c with 1/n cells (u = -uvw_u pixels): a random set plus
    rows whose conjugate footprints reach past the grid edge (u in [-29.5,
    -28.5] puts the conjugate centre at 61..62 of 64)."""
    rng = np.random.default_rng(21)
    uvw, freq, cell, vis, w, flag, n_uv = _random_problem(rng, nt=2, nb=40, nchan=1, n=n)
    cell = (1.0 / n, 1.0 / n)
    freq = np.array([C])
    uvw[..., :2] = rng.uniform(-28.0, 28.0, size=uvw[..., :2].shape)
    edge = np.array([29.0, 29.3, 28.6, 29.5, -28.4, 0.5])
    uvw[0, :edge.size, 0] = edge
    uvw[1, :edge.size, 1] = edge
    return uvw, freq, cell, w[:, :, :1], (n, n)


@pytest.mark.parametrize("n", [64, 63])
def test_grid_imaging_weight_matches_brute_force(oracle, n):
    """do_imaging_weight (the conjugate scatter of make_gridding_convolution_function's
    PSF): the oracle against the cell-centric sum, on an even and an odd grid,
    with conjugate footprints that reach past the grid edge."""
    uvw, freq, cell, w, n_uv = imaging_weight_edge_problem(n)
    g, sw = oracle.standard_grid(None, uvw, w, freq, n_uv, cell, do_psf=True,
                                 do_imaging_weight=True, complex_grid=False)
    gb, swb = brute_force_imaging_weight_psf(oracle, uvw, w, freq, cell, n_uv)
    assert np.max(np.abs(g - gb)) <= 1e-12 * np.max(np.abs(gb))
    np.testing.assert_allclose(sw, swb, rtol=1e-12)
    # the conjugate half is a second copy of every footprint: the sum of
    # weights doubles that of the plain PSF of the averaged weights, and the
    # grid sum falls short of twice the plain one by the clipped cells only
    wa = np.broadcast_to(((w[..., 0] + w[..., 1]) / 2.0)[..., None], w.shape).copy()
    g1, sw1 = oracle.standard_grid(None, uvw, wa, freq, n_uv, cell, do_psf=True,
                                   complex_grid=False)
    np.testing.assert_allclose(sw, 2 * sw1, rtol=1e-12)
    clipped = 2 * g1.sum() - g.sum()
    if n == 64:
        assert clipped > 1e-3 * g1.sum() / (w.shape[0] * w.shape[1])  # some cells fell off
    else:
        assert abs(clipped) <= 1e-12 * g1.sum()


def test_grid_invariants(oracle):
    rng = np.random.default_rng(12)
    uvw, freq, cell, vis, w, flag, n_uv = _random_problem(rng, nt=4, nb=60)
    g, sw = oracle.standard_grid(vis, uvw, w, freq, n_uv, cell)
    # norm of every on-grid sample lies in [5.3427, 5.3441] (support 7, ov 100)
    norms, total = [], np.zeros(2, dtype=np.complex128)
    for t in range(uvw.shape[0]):
        for b in range(uvw.shape[1]):
            for c in range(freq.size):
                gs, s1 = oracle.standard_grid(vis[t:t + 1, b:b + 1, c:c + 1], uvw[t:t + 1, b:b + 1],
                                              w[t:t + 1, b:b + 1, c:c + 1], freq[c:c + 1], n_uv,
                                              cell)
                if s1[0, 0] == 0:
                    continue
                nk = s1[0, 0] / w[t, b, c, 0]
                norms.append(nk)
                total += vis[t, b, c, :2] * w[t, b, c, :2] * nk
    norms = np.array(norms)
    assert norms.size > 50
    # norm = S(ou) * S(ov) with S(o) = sum_i cgk[|100 i + o|], i = -3..3; a
    # dropped or doubled tap moves it out of [min S^2, max S^2]
    cgk = oracle.create_prolate_spheroidal_kernel_1D(100, 7)
    S = np.array([sum(cgk[abs(100 * i + o)] for i in range(-3, 4)) for o in range(-50, 51)])
    assert norms.min() >= S.min() ** 2 * (1 - 1e-14)
    assert norms.max() <= S.max() ** 2 * (1 + 1e-14)
    assert 5.34 < S.min() ** 2 < S.max() ** 2 < 5.35
    np.testing.assert_allclose(g.sum(axis=(2, 3))[0], total, rtol=1e-12)


def test_grid_pixel_centre_footprint(oracle):
    # f = c, dl = 1/N: uvw = (-5, 3) lands exactly on pixel (37, 29) of a 64 grid
    n = 64
    g, sw = oracle.standard_grid(np.ones((1, 1, 1, 1), complex), np.array([[[-5.0, 3.0, 0]]]),
                                 np.ones((1, 1, 1, 1)), np.array([C]), (n, n),
                                 (1.0 / n, 1.0 / n))
    cgk = oracle.create_prolate_spheroidal_kernel_1D(100, 7)
    k1 = cgk[[200, 100, 0, 100, 200]]
    nz = np.argwhere(g[0, 0] != 0)
    assert len(nz) == 25
    np.testing.assert_array_equal(g[0, 0, 35:40, 27:32].real, np.outer(k1, k1))
    assert abs(sw[0, 0] - k1.sum() ** 2) < 1e-14


def test_grid_flags_nan_and_offgrid(oracle):
    rng = np.random.default_rng(13)
    uvw, freq, cell, vis, w, flag, n_uv = _random_problem(rng, frac_flag=0.3)
    img, g, sw = oracle.grid2image(vis, uvw, w, flag, freq, n_uv, cell)
    w0 = w.copy()
    w0[flag] = 0
    g2, sw2 = oracle.standard_grid(vis, uvw, w0, freq, n_uv, cell)
    np.testing.assert_array_equal(g, g2)
    uvw_nan = uvw.copy()
    uvw_nan[0, :5, 0] = np.nan
    _, g3, _ = oracle.grid2image(vis, uvw_nan, w, flag, freq, n_uv, cell)
    w1 = w0.copy()
    w1[0, :5] = 0
    g4, _ = oracle.standard_grid(vis, uvw, w1, freq, n_uv, cell)
    np.testing.assert_allclose(g3, g4, rtol=0, atol=1e-12)
    # far off-grid samples contribute nothing
    g5, sw5 = oracle.standard_grid(vis, uvw * 1e6, w, freq, n_uv, cell)
    assert not g5.any() and not sw5.any()


def test_reference_known_answer_point_sources(oracle):
    """test_standard_grid.py:14-111: 200x200 regular uv grid of
    conj(FFT(model)) visibilities; dirty peaks and PSF centre."""
    with open(GOLDEN) as f:
        gold = json.load(f)["standard_grid_point_sources"]
    nant, maxuv = gold["nant"], gold["maxUV"]
    freq = np.array([gold["freq"]])
    cell = 0.5 * C / freq[0] / maxuv
    mod = np.zeros((nant, nant))
    for (x, y, s) in gold["sources"]:
        mod[x, y] = s
    ft = np.conj(oracle.fft_lm_to_uv(mod, axes=[0, 1]))
    ax = np.linspace(-maxuv, maxuv, nant)
    uu, vv = np.meshgrid(ax, ax, indexing="ij")
    uvw = np.stack([uu.ravel(), vv.ravel(), np.zeros(nant * nant)], axis=1)[None]
    vis = ft.reshape(1, nant * nant, 1, 1)
    w = np.ones(vis.shape)
    gs = (nant, nant)
    g, sw = oracle.standard_grid(vis, uvw, w, freq, gs, (cell, cell))
    _, corr = oracle.create_prolate_spheroidal_kernel(100, 7, gs)
    d = oracle.ifft_uv_to_lm(g, axes=[2, 3]) / corr * nant * nant / sw
    for (x, y, thr) in gold["dirty_thresholds"]:
        assert d[0, 0, x, y] > thr
    g, sw = oracle.standard_grid(vis, uvw, w, freq, gs, (cell, cell), do_psf=True)
    p = oracle.ifft_uv_to_lm(g, axes=[2, 3]) / corr * nant * nant / sw
    assert abs(p[0, 0, nant // 2, nant // 2] - 1.0) < gold["psf_centre_tol"]


def test_cube_constant_visibilities_closed_form(oracle):
    """test_standard_grid.py:113-132 on synthetic uvw: channel k carries the
    constant k+1, so dirty[k, :, c, c] = (k+1) / grdsf(0)^2."""
    rng = np.random.default_rng(14)
    nt, nb, nchan, npol, n = 5, 50, 4, 2, 100
    uvw = rng.uniform(-900, 900, size=(nt, nb, 3))
    freq = np.linspace(1.0e9, 1.1e9, nchan)
    cell = 0.5 * C / freq[-1] / 1400.0
    vis = np.ones((nt, nb, nchan, npol), complex) * (np.arange(nchan) + 1)[None, None, :, None]
    img, _, _ = oracle.grid2image(vis, uvw, np.ones(vis.shape), None, freq, (n, n),
                                  (-cell, cell), chan_mode="cube")
    g0 = 0.9999996674 ** 2
    for k in range(nchan):
        assert abs(img[k, 0, n // 2, n // 2] - (k + 1) / g0) < 1e-9 * (k + 1)


# ---------------------------------------------------------------- degridding
def test_degrid_constant_grid_is_exact(oracle):
    """test_standard_degrid.py:122-240: a constant grid of 2 degrids to
    exactly 2.0."""
    rng = np.random.default_rng(15)
    npix, maxuv = 200, 2e3
    cell = 0.5 * C / 1.4e9 / maxuv
    grid = np.full((1, 1, npix, npix), 2.0 + 0j)
    uvw = np.zeros((1, 4096, 3))
    uvw[0, :, :2] = rng.uniform(-maxuv, maxuv, size=(4096, 2))
    vals = np.zeros((1, 4096, 1, 1), complex)
    oracle.dgrid(uvw, vals, np.zeros(vals.shape, bool), np.array([cell, cell, 0]),
                 np.array([npix // 2, npix // 2, 0]), grid, np.array([1.4e9]), 7, 100,
                 oracle.create_prolate_spheroidal_kernel_1D(100, 7), np.zeros(1, int),
                 np.zeros(1, int))
    assert np.max(np.abs(vals)) == 2.0


def test_degrid_is_normalised_adjoint_of_grid(oracle):
    """For one sample at a non-tie position, gridding (weight 1) produces its
    footprint C_k and sum norm_k; degridding any grid g gives <C_k, g>/norm_k.
    A transposed axis or a sign error in either kernel breaks this."""
    rng = np.random.default_rng(16)
    n = 48
    cell = (-0.5 * C / 1.2e9 / 1500.0, 0.5 * C / 1.2e9 / 1500.0)
    g = rng.standard_normal((1, 1, n, n)) + 1j * rng.standard_normal((1, 1, n, n))
    for _ in range(40):
        uvw = np.zeros((1, 1, 3))
        uvw[0, 0, :2] = rng.uniform(-1000, 1000, size=2)
        gk, sk = oracle.standard_grid(np.ones((1, 1, 1, 1), complex), uvw, np.ones((1, 1, 1, 1)),
                                      np.array([1.2e9]), (n, n), cell)
        m = oracle.degrid(uvw, np.zeros((1, 1, 1, 1), bool), np.array([1.2e9]), g, cell)
        if sk[0, 0] == 0:
            assert m[0, 0, 0, 0] == 0
            continue
        expect = np.sum(gk[0, 0] * g[0, 0]) / sk[0, 0]
        assert abs(m[0, 0, 0, 0] - expect) < 1e-12 * abs(expect) + 1e-14


def test_degrid_flags_and_nan_rows(oracle):
    rng = np.random.default_rng(17)
    uvw, freq, cell, vis, w, flag, n_uv = _random_problem(rng, frac_flag=0.2)
    grid = rng.standard_normal((1, 2) + n_uv) + 0j
    m = oracle.degrid(uvw, flag, freq, grid, cell)
    assert np.all(m[flag] == 0)
    uvw2 = uvw.copy()
    uvw2[1, 3, 1] = np.nan
    m2 = oracle.degrid(uvw2, flag, freq, grid, cell)
    assert np.all(m2[1, 3] == 0)
    mask = np.ones(m.shape, bool)
    mask[1, 3] = False
    np.testing.assert_array_equal(m2[mask], m[mask])


def test_degrid_skips_zero_norm_and_foreign_pols(oracle):
    """dgrid_numba adds nvalue / norm only when norm != 0
    (standard_degrid.py:482-485) and skips a polarisation whose polmap entry
    is outside the grid (:419-420): with a kernel table that is zero on the
    samples' footprints the values stay as they were (no NaN), and a pol
    mapped to plane 5 of a 1-plane grid is never written.  The second pol
    (mapped in range) matches the brute-force sum through the same table."""
    rng = np.random.default_rng(23)
    uvw, freq, cell, vis, w, flag, n_uv = _random_problem(rng)
    grid = rng.standard_normal((1, 1) + n_uv) + 1j * rng.standard_normal((1, 1) + n_uv)
    nt, nb, nf = uvw.shape[0], uvw.shape[1], freq.size
    sc = np.array([-cell[0] * n_uv[0], -cell[1] * n_uv[1], 0.0])
    of = np.array([n_uv[0] / 2, n_uv[1] / 2, 0.0])
    start = np.full((nt, nb, nf, 2), 0.25 - 0.5j)
    vals = start.copy()
    zero = np.zeros(400)
    oracle.dgrid(uvw, vals, np.zeros(vals.shape, bool), sc, of, grid, freq, 7, 100, zero,
                 np.zeros(nf, np.int64), np.array([0, 0], np.int64))
    np.testing.assert_array_equal(vals, start)
    vals = np.zeros((nt, nb, nf, 2), complex)
    conv = oracle.create_prolate_spheroidal_kernel_1D(100, 7)
    oracle.dgrid(uvw, vals, np.zeros(vals.shape, bool), sc, of, grid, freq, 7, 100, conv,
                 np.zeros(nf, np.int64), np.array([5, 0], np.int64))
    assert np.all(vals[..., 0] == 0)
    ref = oracle.degrid(uvw, np.zeros((nt, nb, nf, 1), bool), freq, grid, cell)
    np.testing.assert_array_equal(vals[..., 1], ref[..., 0])


def test_fft_known_answers(oracle):
    """test_fft_ifft.py:9-24 and test_standard_image_grid_preparation.py:127-142."""
    sky = np.zeros((128, 128))
    sky[64, 64] = 1
    sky[50, 40] = 1
    rt = oracle.ifft_uv_to_lm(oracle.fft_lm_to_uv(sky, (0, 1)), (0, 1))
    assert np.max(np.abs(sky - rt)) < 1e-10
    ones = np.ones((200, 200))
    uv = oracle.fft_lm_to_uv(ones, (0, 1))
    assert uv[100, 100] == 200 * 200
    vis = np.zeros((200, 200), complex)
    vis[100, 100] = 10
    assert abs(oracle.ifft_uv_to_lm(vis, (0, 1)).max() - 10.0 / 200 / 200) < 1e-15


# ------------------------------------------------------- measurement operator
def test_oracle_grid_adjoint_is_transpose_of_degrid(oracle):
    rng = np.random.default_rng(18)
    n = 40
    cell = (-0.5 * C / 1.2e9 / 1500.0, 0.5 * C / 1.2e9 / 1500.0)
    uvw = np.zeros((2, 60, 3))
    uvw[..., :2] = rng.uniform(-1200, 1200, size=(2, 60, 2))
    freq = np.array([1.0e9, 1.2e9])
    g = rng.standard_normal((1, 2, n, n)) + 1j * rng.standard_normal((1, 2, n, n))
    y = rng.standard_normal((2, 60, 2, 2)) + 1j * rng.standard_normal((2, 60, 2, 2))
    flag = rng.random(y.shape) < 0.2
    dg = oracle.degrid(uvw, flag, freq, g, cell)
    dhy = oracle.grid_adjoint(uvw, y, flag, freq, (n, n), cell, 1, 2)
    lhs = np.vdot(y, dg)
    rhs = np.vdot(dhy, g)
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(dg) * np.linalg.norm(y)


def test_oracle_measurement_operator_adjoint(oracle):
    rng = np.random.default_rng(19)
    n = 32
    cell = (-0.5 * C / 1.2e9 / 1200.0, 0.5 * C / 1.2e9 / 1200.0)
    uvw = np.zeros((1, 200, 3))
    uvw[..., :2] = rng.uniform(-1000, 1000, size=(1, 200, 2))
    freq = np.array([1.2e9])
    x = rng.standard_normal((1, 1, n, n)) + 1j * rng.standard_normal((1, 1, n, n))
    y = rng.standard_normal((1, 200, 1, 1)) + 1j * rng.standard_normal((1, 200, 1, 1))
    ax = oracle.measure(x, uvw, freq, cell)
    ahy = oracle.measure_adjoint(y, uvw, freq, (n, n), cell)
    assert abs(np.vdot(y, ax) - np.vdot(ahy, x)) <= 1e-12 * np.linalg.norm(ax) * np.linalg.norm(y)
