"""av_grid_noalloc (binning + gridding with no device allocation, its plan
in the library's static scratch array) against av_plan_create +
av_plan_set_uvw + av_grid on the same inputs, through the C ABI: two calls
in a row, the second with fewer samples and a smaller grid, so a stale
plan left in the scratch array by the first call would show.  Also: a call
whose plan exceeds the scratch array is refused with a clear error."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

C = 299792458.0


def rel(a, b):
    den = np.max(np.abs(b))
    return float(np.max(np.abs(np.asarray(a) - b)) / (den if den > 0 else 1.0))


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def problem(seed, n_time, n, nchan, npol):
    from astroviper_amd import synth
    arr = synth.ArrayConfig(n_time=n_time, n_chan=nchan, n_pol=npol, seed=seed)
    obs = synth.make_observation(arr, synth.SkyConfig())
    vis, w, flag = synth.host_visibilities(obs, synth.SkyConfig(), flag_frac=0.1,
                                           weight_range=(0.5, 1.5))
    maxuv = np.max(np.hypot(obs.uvw[..., 0], obs.uvw[..., 1]))
    c = 0.45 * C / obs.freq[-1] / maxuv
    return obs, vis, w, flag, (n, n), (-c, c)


def grid_both_ways(gpu, lib, obs, vis, w, flag, n_uv, cell, prec, cube, tile=(16, 16)):
    from astroviper_amd import _lib
    from astroviper_amd.device import ps_kernel_1d
    nt, nb, nchan, npol = w.shape
    rows = nt * nb
    n_img = nchan if cube else 1
    rdt = torch.float32 if prec == 32 else torch.float64
    cdt = torch.complex64 if prec == 32 else torch.complex128
    code = _lib.AV_F32 if prec == 32 else _lib.AV_F64
    uvw = torch.from_numpy(np.ascontiguousarray(obs.uvw.reshape(rows, 3))).to(gpu)
    freq = torch.from_numpy(obs.freq).to(gpu)
    cmap = torch.arange(nchan, dtype=torch.int32, device=gpu) if cube else \
        torch.zeros(nchan, dtype=torch.int32, device=gpu)
    v = torch.from_numpy(vis.reshape(rows * nchan, npol)).to(gpu, cdt)
    wt = torch.from_numpy(w.reshape(rows * nchan, npol)).to(gpu, rdt)
    fl = torch.from_numpy(flag.reshape(rows * nchan, npol).astype(np.uint8)).to(gpu)
    cgk = ps_kernel_1d(100, 7, gpu)
    shape = (n_img, npol) + tuple(n_uv)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    # the allocating path
    g1 = torch.empty(shape, dtype=cdt, device=gpu)
    s1 = torch.empty((n_img, npol), dtype=torch.float64, device=gpu)
    h = ctypes.c_void_p()
    _lib.check(lib.av_plan_create(rows, nchan, n_img, n_uv[0], n_uv[1], cell[0], cell[1], 7, 100,
                                  tile[0], tile[1], 0, 0, ctypes.byref(h)), "av_plan_create")
    _lib.check(lib.av_plan_set_uvw(h, _p(uvw), _p(freq), _p(cmap), None, s), "av_plan_set_uvw")
    _lib.check(lib.av_grid(h, _p(v), _p(wt), _p(fl), npol, _p(cgk), _p(g1), _p(s1), code, 0, s),
               "av_grid")
    # the allocation-free path, into buffers holding garbage
    g2 = torch.full(shape, 7.0 + 3.0j, dtype=cdt, device=gpu)
    s2 = torch.full((n_img, npol), -5.0, dtype=torch.float64, device=gpu)
    _lib.check(lib.av_grid_noalloc(_p(uvw), rows, _p(freq), nchan, _p(cmap), n_img, n_uv[0],
                                   n_uv[1], cell[0], cell[1], 7, 100, tile[0], tile[1], _p(v),
                                   _p(wt), _p(fl), npol, _p(cgk), _p(g2), _p(s2), code, 0, s),
               "av_grid_noalloc")
    torch.cuda.synchronize()
    _lib.check(lib.av_plan_destroy(h), "av_plan_destroy")
    return g1.cpu().numpy(), s1.cpu().numpy(), g2.cpu().numpy(), s2.cpu().numpy()


@pytest.mark.parametrize("prec", [64, 32])
def test_noalloc_matches_plan_path(gpu, oracle, prec):
    from astroviper_amd import _lib
    lib = _lib.lib()
    # first call: 27 antennas x 60 times x 4 channels x 2 pols on 256^2, cube
    # second call: 27 antennas x 20 times x 2 channels x 4 pols on 128^2, continuum
    for seed, n_time, n, nchan, npol, cube in ((1, 60, 256, 4, 2, True),
                                                (2, 20, 128, 2, 4, False)):
        obs, vis, w, flag, n_uv, cell = problem(seed, n_time, n, nchan, npol)
        if prec == 32:
            vis = vis.astype(np.complex64).astype(np.complex128)
            w = w.astype(np.float32).astype(np.float64)
        g1, s1, g2, s2 = grid_both_ways(gpu, lib, obs, vis, w, flag, n_uv, cell, prec, cube)
        # record order inside a bin comes from atomic cursors, so sums may
        # differ in their last bits between any two runs of either path: in
        # FP32 a continuum bin here sums thousands of records, and two orders
        # of one sum differ by up to ~1e-6 of the grid maximum (measured
        # 1.04e-6); a stale or wrong plan would differ by O(1)
        tol = 1e-12 if prec == 64 else 1e-5
        assert rel(g2, g1) < tol, (seed, prec)
        assert rel(s2, s1) < 1e-14, (seed, prec)
        # and both are the oracle's
        w0 = np.where(flag, 0.0, w)
        gref, swref = oracle.standard_grid(vis, obs.uvw, w0, obs.freq, n_uv, cell,
                                           chan_mode="cube" if cube else "continuum")
        assert rel(g2, gref) < (1e-12 if prec == 64 else 1e-4)
        assert rel(s2, swref) < 1e-12


def test_noalloc_refuses_plans_beyond_scratch(gpu):
    """A plan larger than the 32 MiB scratch array is refused before any
    work is queued."""
    from astroviper_amd import _lib
    lib = _lib.lib()
    rows, nchan, n = 2_000_000, 4, 256
    uvw = torch.zeros((rows, 3), dtype=torch.float64, device=gpu)
    freq = torch.full((nchan,), 1.0e9, dtype=torch.float64, device=gpu)
    v = torch.zeros((rows * nchan, 1), dtype=torch.complex64, device=gpu)
    wt = torch.ones((rows * nchan, 1), dtype=torch.float32, device=gpu)
    g = torch.zeros((1, 1, n, n), dtype=torch.complex64, device=gpu)
    sw = torch.zeros((1, 1), dtype=torch.float64, device=gpu)
    from astroviper_amd.device import ps_kernel_1d
    cgk = ps_kernel_1d(100, 7, gpu)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.av_grid_noalloc(_p(uvw), rows, _p(freq), nchan, None, 1, n, n, -1e-5, 1e-5, 7, 100,
                             16, 16, _p(v), _p(wt), None, 1, _p(cgk), _p(g), _p(sw), _lib.AV_F32,
                             0, s)
    assert rc == _lib.AV_ERR_INVALID
    assert b"static scratch" in lib.av_last_error()
