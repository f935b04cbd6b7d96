"""Benchmark: visibilities/s of the batched FP32 standard gridder (grid +
degrid, support 7, oversampling 100) on one MI355X.

Workload (default): 64 antennas (2016 baselines) x 1000 time samples x 64
channels x 4 polarisations = 516,096,000 complex64 visibilities with FP64
uvw, 10% flagged, weights U(0.5, 1.5); 2048 x 2048 grid, 1" cells, cube
mode (256 planes).  Inputs are generated on the GPU and stay resident.

A step is one pass of the hot path over the batch: bin the samples
(av_plan_set_uvw), grid them (av_grid) and degrid them back (av_degrid).

    python bench.py [--steps K] [--warmup W] [--mode cube|continuum]

Prints one JSON line.  The CPU baseline is the FP64 C oracle (a loop-by-loop
restatement of the reference's numba kernels) timed single-threaded on a
bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def measured_traffic(kernel, visibilities, n_uv, mode, tile):
    """HBM bytes per launch of `kernel` on this workload, as measured with
    rocprofv3 counters (FETCH_SIZE + WRITE_SIZE, each scaled by the ratio of
    a copy of known size to what the same counter reads for it) and recorded
    in profiles/traffic.json; None when no measurement of this exact
    workload and tile is recorded there."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    for row in table.get("measurements", []):
        if (row.get("kernel") == kernel and row.get("visibilities") == visibilities
                and row.get("n_uv") == list(n_uv) and row.get("mode") == mode
                and row.get("tile") == list(tile)):
            return row.get("bytes_per_launch")
    return None


def host_threads():
    """Host threads this process may use: OMP_NUM_THREADS when set (the GPU
    hosts set it to their CPU share), else the CPUs it may run on, at most 16."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(16, os.cpu_count() or 1))


def _oracle_pair(O, uvw, vis, w, fl, freq, n_uv, cell, mode):
    """One FP64 oracle grid + degrid of the given sample; (t_grid, t_degrid)."""
    t0 = time.perf_counter()
    g, _ = O.standard_grid(vis, uvw, w, freq, n_uv, cell, chan_mode=mode)
    t1 = time.perf_counter()
    O.degrid(uvw, fl, freq, g, cell)
    return t1 - t0, time.perf_counter() - t1


def cpu_baseline(obs, uvw_d, vis_d, w_d, flag_d, n_uv, cell, cube, leg_s=3.0, reps=3):
    """The FP64 C oracle (a loop-by-loop restatement of the reference's numba
    kernels) timed on the host: grid + degrid of the first rows (time-major)
    of the benchmark workload, the median of `reps` repetitions.

    * one core: the first 4 channels, rows grown until one pair takes about
      `leg_s` seconds (the reference's numba kernels are single-threaded);
    * all host threads: one channel per thread (each its own cube plane, as
      an OpenMP loop over planes would run), the same number of rows, the
      channels' pairs run concurrently (the oracle's C calls release the
      interpreter lock).
    Per-visibility cost does not depend on the channel."""
    import concurrent.futures as cf

    from oracle import oracle as O

    nt, nb, _ = obs.uvw.shape
    npol = vis_d.shape[2]
    nchan = vis_d.shape[1]
    total_rows = nt * nb
    mode = "cube" if cube else "continuum"

    def sample(rows, c0, c1):
        sel = slice(0, rows)
        nch = c1 - c0
        uvw = uvw_d[sel].cpu().numpy().reshape(1, rows, 3)
        vis = vis_d[sel, c0:c1].cpu().numpy().astype(np.complex128).reshape(1, rows, nch, npol)
        w = w_d[sel, c0:c1].cpu().numpy().astype(np.float64).reshape(1, rows, nch, npol)
        fl = flag_d[sel, c0:c1].cpu().numpy().astype(bool).reshape(1, rows, nch, npol)
        w[fl] = 0
        return uvw, vis, w, fl, obs.freq[c0:c1]

    # one core: size the sample, then time it
    nch1 = min(4, nchan)
    rows = 20000
    while True:
        rows = min(rows, total_rows)
        smp = sample(rows, 0, nch1)
        tg, td = _oracle_pair(O, *smp[:4], smp[4], n_uv, cell, mode)
        if tg + td > leg_s / 4 or rows * 4 > total_rows:
            break
        rows *= 4
    rows = min(total_rows, int(rows * max(1.0, leg_s / max(tg + td, 1e-3))))
    smp = sample(rows, 0, nch1)
    times1 = [sum(_oracle_pair(O, *smp[:4], smp[4], n_uv, cell, mode)) for _ in range(reps)]
    t1 = float(np.median(times1))
    nvis1 = rows * nch1 * npol
    del smp

    # all threads: one channel each, the same rows
    nthr = min(host_threads(), nchan)
    chans = [sample(rows, c, c + 1) for c in range(nthr)]
    with cf.ThreadPoolExecutor(max_workers=nthr) as ex:
        def run_all():
            t0 = time.perf_counter()
            list(ex.map(lambda smp: _oracle_pair(O, *smp[:4], smp[4], n_uv, cell, mode), chans))
            return time.perf_counter() - t0
        timesn = [run_all() for _ in range(reps)]
    tn = float(np.median(timesn))
    nvisn = rows * nthr * npol
    del chans
    model = cpu_model()
    return {
        "value": round(nvisn / tn / 1e6, 3),
        "unit": "M visibilities/sec",
        "cores": nthr,
        "kind": "port",
        "sample": f"first {rows} of {total_rows} rows x {nthr} of {nchan} channels x {npol} pols "
                  f"= {nvisn} visibilities, FP64 grid + degrid, one channel (cube plane) per "
                  f"thread on {nthr} threads, median of {reps}: {tn:.2f} s",
        "single_core": {"value": round(nvis1 / t1 / 1e6, 3), "unit": "M visibilities/sec",
                        "sample": f"first {rows} rows x first {nch1} channels x {npol} pols = "
                                  f"{nvis1} visibilities, median of {reps}: {t1:.2f} s"},
        "nproc": os.cpu_count(),
        "threads_allowed": host_threads(),
        "cpu_model": model,
    }


def small_workload_leg(dev, reps=10):
    """The 27-antenna workload (100 times x 4 channels x 2 pols = 280,800
    visibilities, 256 x 256, 5" cells, cube) in full: the FP64 oracle on one
    host core against the batched FP32 GPU path (plan + grid + degrid, HIP
    events, median of `reps`), and the GPU results' errors against the
    oracle: grid, dirty image, PSF, sum of weights, model visibilities."""
    import torch

    from astroviper_amd import synth
    from astroviper_amd.device import GridPlan, dirty_image
    from oracle import oracle as O

    cfg = synth.CONFIG1
    obs = synth.make_observation(cfg["array"], cfg["sky"])
    vis, w, flag = synth.host_visibilities(obs, cfg["sky"], flag_frac=0.1, weight_range=(0.5, 1.5))
    vis = vis.astype(np.complex64).astype(np.complex128)
    w = w.astype(np.float32).astype(np.float64)
    c = cfg["cell_arcsec"] * synth.ARCSEC
    cell, n_uv = (-c, c), cfg["n_uv"]
    nt, nb, nchan, npol = w.shape
    rows, V = nt * nb, w.size
    w0 = np.where(flag, 0.0, w)
    tcpu = []
    for _ in range(3):
        t0 = time.perf_counter()
        gref, swref = O.standard_grid(vis, obs.uvw, w0, obs.freq, n_uv, cell, chan_mode="cube")
        mref = O.degrid(obs.uvw, flag, obs.freq, gref, cell)
        tcpu.append(time.perf_counter() - t0)
    pref, pswref = O.standard_grid(None, obs.uvw, w0, obs.freq, n_uv, cell, chan_mode="cube",
                                   do_psf=True)
    iref, psfref = O.normalise(gref, swref), O.normalise(pref, pswref)

    plan = GridPlan(rows, nchan, nchan, n_uv, cell, 7, 100, tile=(16, 16), device=dev,
                    n_baseline=nb)
    uvw_t = torch.from_numpy(obs.uvw.reshape(rows, 3)).to(dev)
    f_t = torch.from_numpy(obs.freq).to(dev)
    v = torch.from_numpy(vis.reshape(rows * nchan, npol)).to(dev, torch.complex64)
    wt = torch.from_numpy(w.reshape(rows * nchan, npol)).to(dev, torch.float32)
    fl = torch.from_numpy(flag.reshape(rows * nchan, npol).astype(np.uint8)).to(dev)
    g = torch.empty((nchan, npol) + tuple(n_uv), dtype=torch.complex64, device=dev)
    sw = torch.empty((nchan, npol), dtype=torch.float64, device=dev)
    m = torch.empty((rows * nchan, npol), dtype=torch.complex64, device=dev)

    def step():
        plan.set_uvw(uvw_t, f_t)
        plan.grid(v, wt, fl, grid=g, sumwt=sw)
        plan.degrid(g, fl, m)

    step()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        step()
        e1.record()
        e1.synchronize()
        ts.append(e0.elapsed_time(e1))
    tgpu = float(np.median(ts))
    img = dirty_image(g, sw).cpu().numpy()
    pg, psw = plan.grid(None, wt, fl, psf=True)
    psf = dirty_image(pg, psw).cpu().numpy()
    _, corr = O.create_prolate_spheroidal_kernel(100, 7, np.array(n_uv))
    inner = corr >= 1e-2

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    def img_err(a, b):
        d = np.abs(a - b)
        return {"full": float(d.max() / np.abs(b).max()),
                "taper_ge_1e-2": float(d[..., inner].max() / np.abs(b).max())}

    # the GPU degrids its own FP32 grid; the oracle model is from the FP64 grid
    mm = m.cpu().numpy().reshape(mref.shape)
    plan.close()
    drop_in = drop_in_leg(obs, vis, w, flag, cell, n_uv, gref, iref, mref)
    return {"workload": f"27 antennas x {nt} times x {nchan} chan x {npol} pol = {V} vis, "
                        f"{n_uv[0]}x{n_uv[1]} cube, 5\" cells, 10% flagged, weights U(0.5,1.5)",
            "cpu_fp64_one_core": {"value": round(V / float(np.median(tcpu)) / 1e6, 3),
                                  "unit": "M visibilities/sec", "median_s": float(np.median(tcpu))},
            "gpu_fp32": {"value": round(V / (tgpu / 1e3) / 1e6, 1), "unit": "M visibilities/sec",
                         "ms": round(tgpu, 4), "what": "plan + grid + degrid, median of events"},
            "max_rel_err_vs_fp64_oracle": {
                "grid": rel(g.cpu().numpy(), gref), "sumwt": rel(sw.cpu().numpy(), swref),
                "dirty": img_err(img, iref), "psf": img_err(psf, psfref),
                "degrid_from_own_grid": rel(mm, mref)},
            "drop_in": drop_in}


def drop_in_leg(obs, vis, w, flag, cell, n_uv, gref, iref, mref, reps=5):
    """The reference entry points end to end on the 27-antenna workload:
    grid2image_spheroid_ms4 (standard_grid.py:607-708) and
    degrid_spheroid_ms4 (standard_degrid.py:488-608) from an MSv4-shaped
    dataset of host arrays to host results (copies to the device, a plan per
    call, FP64 gridding and FFT as the complex128 data select, copies back),
    median of `reps` after one warm-up call, beside the FP64 oracle doing the
    same work on one host core (gridding + normalisation, degridding)."""
    from astroviper_amd.msv4 import make_visibility_dataset
    from astroviper_amd.standard_degrid import degrid_spheroid_ms4
    from astroviper_amd.standard_grid import grid2image_spheroid_ms4
    from oracle import oracle as O

    nt, nb, nchan, npol = w.shape
    pix = np.array(cell)

    def ds():
        return make_visibility_dataset(vis.copy(), obs.uvw.copy(), w.copy(), flag.copy(), obs.freq)

    def t_grid():
        d = ds()
        img = np.zeros((nchan, npol) + tuple(n_uv))
        t0 = time.perf_counter()
        grid2image_spheroid_ms4(d, img, pix, 7, 100, dopsf=False, column="VISIBILITY",
                                chan_mode="cube")
        return time.perf_counter() - t0, img

    def t_degrid():
        d = ds()
        t0 = time.perf_counter()
        degrid_spheroid_ms4(d, gref, pix, 7, 100)
        return time.perf_counter() - t0, np.asarray(d["VISIBILITY_MODEL"].data)

    t_grid()
    t_degrid()
    tg, img = zip(*[t_grid() for _ in range(reps)])
    td, mod = zip(*[t_degrid() for _ in range(reps)])
    w0 = np.where(flag, 0.0, w)
    tcg, tcd = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        g, s = O.standard_grid(vis, obs.uvw, w0, obs.freq, n_uv, cell, chan_mode="cube")
        O.normalise(g, s)
        tcg.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        O.degrid(obs.uvw, flag, obs.freq, gref, cell)
        tcd.append(time.perf_counter() - t0)

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    return {"what": "host dataset in, host image / model visibilities out: copies, binning, "
                    "FP64 gridding or degridding, FFT (gridding), one plan per call",
            "grid2image_spheroid_ms4_ms": round(1e3 * float(np.median(tg)), 3),
            "degrid_spheroid_ms4_ms": round(1e3 * float(np.median(td)), 3),
            "cpu_fp64_one_core_grid_image_ms": round(1e3 * float(np.median(tcg)), 3),
            "cpu_fp64_one_core_degrid_ms": round(1e3 * float(np.median(tcd)), 3),
            "max_rel_err_vs_fp64_oracle": {"dirty": rel(img[-1], iref),
                                           "model": rel(mod[-1].reshape(mref.shape), mref)}}


# --------------------------------------------------------------------------
# Rows beside the hot path: Hogbom CLEAN, imaging weights, the mosaic gridder
# and resident major cycles, each on one GPU with a bounded CPU leg and the
# GPU result compared with the CPU one on the same sample.
#
#     python bench.py --rows hogbom,weights,mosaic,gcf,major_cycle
#
# prints one JSON line per row (the headline line is not printed then).
# --------------------------------------------------------------------------
def _events_ms(fn, reps=3, warm=1):
    """Median HIP-event time (ms, torch's current stream, on which the
    library launches) of fn()."""
    import torch
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        ts.append(a.elapsed_time(b))
    return float(np.median(ts))


def _roof(nbytes, ms, kernel):
    gbs = nbytes / (ms / 1e3) / 1e9
    return {"bound": "hbm", "kernel": kernel, "achieved": round(gbs, 1),
            "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(gbs * 1e9 / HBM_PEAK, 4),
            "algorithmic_bytes": int(nbytes), "avg_ms": round(ms, 4)}


def _ref_hogbom():
    """The reference's own Hogbom extension built from its sources into
    oracle/_ref (None when that build is absent)."""
    ref = os.path.join(ROOT, "oracle", "_ref")
    if ref not in sys.path:
        sys.path.insert(0, ref)
    try:
        import _hogbom_ext
        return _hogbom_ext
    except ImportError:
        return None


def row_hogbom(dev, n=2048, niter=2000, half_box=32, cpu_s=10.0):
    """Hogbom CLEAN of one 2048 x 2048 float32 plane (hclean.cpp:106-210):
    a PSF with a Gaussian main lobe and ring sidelobes over the whole image,
    40 point sources within 32 pixels of the centre, a clean box of the
    central 64 x 64 pixels, gain 0.1, threshold 0, 2000 iterations.

    An iteration subtracts gain * peak * PSF over the window where the
    shifted PSF overlaps the image, W = (n - |dx|)(n - |dy|) pixels for a peak
    (dx, dy) from the centre, and finds the next peak in the box (inside the
    window): 12 bytes per window pixel (residual read and written, PSF
    read).  The peaks lie in the box, so W >= (n - 32)^2 = 0.969 n^2; the
    algorithmic bytes use that bound, so `achieved` is a lower bound."""
    import torch

    from astroviper_amd.deconvolvers import hogbom as H

    rng = np.random.default_rng(11)
    c = n // 2
    yy, xx = np.mgrid[0:n, 0:n]
    r = np.hypot(xx - c, yy - c)
    psf = np.exp(-r ** 2 / 18.0) + 0.05 * np.cos(2 * np.pi * r / 12.0) * np.exp(-r / 200.0)
    psf /= psf[c, c]
    sky = np.zeros((n, n))
    k = 40
    np.add.at(sky, (rng.integers(c - half_box + 1, c + half_box, k),
                    rng.integers(c - half_box + 1, c + half_box, k)), rng.uniform(0.1, 1.0, k))
    dirty = np.real(np.fft.ifft2(np.fft.fft2(sky) * np.fft.fft2(np.fft.ifftshift(psf))))
    dirty = dirty.astype(np.float32)
    psf = psf.astype(np.float32)
    del yy, xx, r
    box = (c - half_box, c + half_box, c - half_box, c + half_box)  # xbeg, xend, ybeg, yend
    d_t = torch.from_numpy(dirty).to(dev)
    p_t = torch.from_numpy(psf).to(dev)
    work = H._work(dev)
    res = {}

    def run(iters):
        rr = d_t.clone()
        mm = torch.zeros_like(rr)
        res["it"] = H.clean_device(rr, p_t, mm, None, box, iters, 0, 0.1, 0.0, work=work)
        res["r"], res["m"] = rr, mm

    ms = _events_ms(lambda: run(niter), reps=3)
    it = res["it"]
    log(f"bench: hogbom {it} iterations in {ms:.2f} ms on the GPU")
    wmin = (n - half_box) ** 2
    out = {"row": "hogbom", "workload": f"Hogbom CLEAN, one {n}x{n} float32 plane, {k} point "
                                         f"sources, clean box {2 * half_box}x{2 * half_box} at "
                                         f"the centre, gain 0.1, {niter} iterations",
           "value": round(it / (ms / 1e3), 1), "unit": "iterations/sec",
           "gpu_ms": round(ms, 3), "iterations": it,
           "us_per_iteration": round(ms * 1e3 / max(it, 1), 3),
           "roofline": _roof(12.0 * wmin * it, ms, "k_hogbom")}
    out["roofline"]["note"] = ("the residual and the PSF (32 MB) fit in the 256 MB "
                               "Infinity Cache, so iterations may run above HBM speed")
    # CPU leg: the reference's own extension (else the oracle's restatement),
    # one core, on the same images, iterations sized to about cpu_s seconds;
    # the GPU run of the same iterations is compared with it
    ext = _ref_hogbom()
    kind = "reference" if ext is not None else "port"
    if ext is None:
        from oracle import oracle as O

        def cpu_clean(iters):
            return O.hogbom_clean(dirty, psf, clean_box=box, max_iter=iters, gain=0.1,
                                  threshold=0.0)
    else:
        def cpu_clean(iters):
            return ext.clean(dirty, psf, clean_box=box, max_iter=iters, gain=0.1, threshold=0.0)
    t0 = time.perf_counter()
    cpu_clean(20)
    per = (time.perf_counter() - t0) / 20
    ci = int(max(20, min(niter, cpu_s / max(per, 1e-6))))
    t0 = time.perf_counter()
    ref = cpu_clean(ci)
    tc = time.perf_counter() - t0
    run(ci)
    dm = float(np.max(np.abs(res["m"].cpu().numpy() - ref["model_image"])))
    dr = float(np.max(np.abs(res["r"].cpu().numpy() - ref["residual_image"])))
    out["cpu_baseline"] = {"value": round(ref["iterations_performed"] / tc, 2),
                           "unit": "iterations/sec", "cores": 1, "kind": kind,
                           "sample": f"the same images, {ref['iterations_performed']} "
                                     f"iterations: {tc:.2f} s"}
    out["parity_vs_cpu"] = {"iterations": ci, "same_iterations": res["it"] == ref[
        "iterations_performed"], "model_max_abs_diff": dm, "residual_max_abs_diff": dr}
    return out


def row_weights(dev, obs, uvw, w, flag, n_uv, cell, robust=0.5, cpu_s=5.0):
    """Briggs imaging weights of the benchmark workload (64 antennas x 1000
    times x 64 channels, 2048 x 2048 cells per channel; calculate_imaging_
    weights.py:28-275): the data weights of polarisations 0 and 3 averaged,
    NaN where flagged; gridded at the nearest pixel and its conjugate
    (grid_imaging_weights_jit, grid_imaging_weights.py:105-206), the Briggs
    factors from the density grid (briggs_weighting.py:1-31), and the
    per-sample imaging weights (degrid_imaging_weights_jit :274-443).

    Algorithmic bytes: grid pass uvw (24 B per row) + weight (8 B per
    sample) + density grid once (8 B per cell); degrid pass the same plus
    the imaging weight written (8 B per sample)."""
    import torch

    from astroviper_amd import _lib
    from astroviper_amd.device import _ptr, _stream

    rows, nchan, npol = w.shape
    S = rows * nchan
    nbl = obs.uvw.shape[1]  # rows are time-major, nbl baselines per time
    w1 = ((w[..., 0].double() + w[..., npol - 1].double()) / 2)
    w1 = torch.where(flag[..., 0].bool(), torch.full_like(w1, float("nan")), w1).contiguous()
    freq = torch.from_numpy(obs.freq).to(dev)
    n_u, n_v = n_uv
    grid = torch.empty((nchan, n_u, n_v), dtype=torch.float64, device=dev)
    sw = torch.empty(nchan, dtype=torch.float64, device=dev)
    iw = torch.empty((rows, nchan), dtype=torch.float64, device=dev)
    nanf = torch.zeros(1, dtype=torch.int32, device=dev)
    bf = torch.empty((2, nchan), dtype=torch.float64, device=dev)

    def grid_w():
        grid.zero_()
        sw.zero_()
        _lib.call("av_grid_imaging_weights", _ptr(uvw), rows, nbl, _ptr(freq), nchan, _ptr(w1), 1,
                  n_u, n_v, float(cell[0]), float(cell[1]), _ptr(grid), _ptr(sw), _stream())

    def factors():
        sq = torch.sum(grid * grid, dim=(1, 2))
        bf[0] = (5.0 * 10.0 ** (-robust)) ** 2 / (sq / sw)
        bf[1] = 1.0

    def degrid_w():
        _lib.call("av_degrid_imaging_weights", _ptr(grid), _ptr(bf), _ptr(uvw), rows, nbl,
                  _ptr(freq), nchan, _ptr(w1), n_u, n_v, float(cell[0]), float(cell[1]),
                  _ptr(iw), _ptr(nanf), _stream())

    t_g = _events_ms(grid_w)
    t_f = _events_ms(factors)
    t_d = _events_ms(degrid_w)
    log(f"bench: weights grid {t_g:.2f} ms, factors {t_f:.2f} ms, degrid {t_d:.2f} ms")
    cells = nchan * n_u * n_v
    b_g = rows * 24 + S * 8 + cells * 8
    b_d = rows * 24 + S * 16 + cells * 8
    tot = t_g + t_f + t_d
    out = {"row": "weights", "workload": f"Briggs (robust {robust}) imaging weights, {rows} rows "
                                          f"x {nchan} channels = {S} samples, {n_u}x{n_v} "
                                          f"density grid per channel",
           "value": round(S / (tot / 1e3) / 1e6, 1), "unit": "M samples/sec",
           "ms": {"grid_imaging_weights": round(t_g, 3), "briggs_factors": round(t_f, 3),
                  "degrid_imaging_weights": round(t_d, 3)},
           "roofline": _roof(b_g, t_g, "k_grid_imaging_weights"),
           "roofline_degrid": _roof(b_d, t_d, "k_degrid_imaging_weights")}
    # CPU leg: the oracle on the first rows x all channels, one core
    from oracle import oracle as O
    nb = obs.uvw.shape[1]
    nt = 4
    while True:
        r_s = nt * nb
        u_h = uvw[:r_s].cpu().numpy().reshape(nt, nb, 3)
        w_h = w1[:r_s].cpu().numpy().reshape(nt, nb, nchan, 1)
        g_h = np.zeros((nchan, 1, n_u, n_v))
        s_h = np.zeros((nchan, 1))
        t0 = time.perf_counter()
        O.grid_imaging_weights(g_h, s_h, u_h, w_h, obs.freq, n_uv, cell)
        sq = np.sum(g_h[:, 0] ** 2, axis=(1, 2))
        bf_h = np.stack([(5.0 * 10.0 ** (-robust)) ** 2 / (sq / s_h[:, 0]), np.ones(nchan)])
        iw_h = O.degrid_imaging_weights(g_h, u_h, w_h, bf_h, obs.freq, n_uv, cell)
        tc = time.perf_counter() - t0
        log(f"bench: weights CPU leg {r_s} rows: {tc:.2f} s")
        if tc > cpu_s / 4 or r_s * 4 > rows:
            break
        nt *= 4
    # the GPU on the same sample
    uvw_s, w1_s = uvw[:r_s].contiguous(), w1[:r_s].contiguous()
    g_s = torch.zeros((nchan, n_u, n_v), dtype=torch.float64, device=dev)
    s_s = torch.zeros(nchan, dtype=torch.float64, device=dev)
    _lib.call("av_grid_imaging_weights", _ptr(uvw_s), r_s, nbl, _ptr(freq), nchan, _ptr(w1_s), 1,
              n_u, n_v, float(cell[0]), float(cell[1]), _ptr(g_s), _ptr(s_s), _stream())
    bf_s = torch.from_numpy(bf_h).to(dev).contiguous()
    iw_s = torch.empty((r_s, nchan), dtype=torch.float64, device=dev)
    _lib.call("av_degrid_imaging_weights", _ptr(g_s), _ptr(bf_s), _ptr(uvw_s), r_s, nbl, _ptr(freq),
              nchan, _ptr(w1_s), n_u, n_v, float(cell[0]), float(cell[1]), _ptr(iw_s),
              _ptr(nanf), _stream())

    def rel(a, b):
        return float(np.nanmax(np.abs(a - b)) / np.nanmax(np.abs(b)))

    iw_ref = np.asarray(iw_h)[..., 0].reshape(r_s, nchan)
    out["cpu_baseline"] = {"value": round(r_s * nchan / tc / 1e6, 3), "unit": "M samples/sec",
                           "cores": 1, "kind": "port",
                           "sample": f"first {r_s} rows x {nchan} channels = {r_s * nchan} "
                                     f"samples, grid + factors + degrid: {tc:.2f} s"}
    iw_g = iw_s.cpu().numpy()
    out["parity_vs_cpu"] = {"samples": r_s * nchan,
                            "density_grid_max_rel": rel(g_s.cpu().numpy(), g_h[:, 0]),
                            "imaging_weight_max_rel": rel(iw_g, iw_ref),
                            "nan_positions_differ": int(np.sum(np.isnan(iw_g) != np.isnan(iw_ref)))}
    return out


def _mosaic_cfs(rng, ws_choices=(7, 9, 11), ov=(8, 8), ncb=3, ncc=4):
    """Complex convolution functions of A-projection shape: a Gaussian
    aperture product with a phase ramp, per (baseline group, channel group),
    one polarisation group; supports per axis from ws_choices."""
    ws = rng.choice(np.asarray(ws_choices), size=(ncb, ncc, 1, 2))
    smax = int(ws.max())
    cu_n, cv_n = ov[0] * (smax + 1) + 1, ov[1] * (smax + 1) + 1
    x = (np.arange(cu_n) - cu_n // 2) / cu_n
    y = (np.arange(cv_n) - cv_n // 2) / cv_n
    cf = np.zeros((ncb, ncc, 1, cu_n, cv_n), dtype=np.complex128)
    for idx in np.ndindex(ncb, ncc, 1):
        a, b_, ph = rng.uniform(4, 8), rng.uniform(4, 8), rng.uniform(-3, 3)
        cf[idx] = (np.exp(-a * x[:, None] ** 2 - b_ * y[None, :] ** 2)
                   * np.exp(1j * ph * (x[:, None] + 0.5 * y[None, :])))
    pg = np.exp(1j * rng.uniform(-0.3, 0.3, size=(1, cu_n, cv_n)))
    return ws, cf, pg, ov


def row_mosaic(dev, n_time=100, cpu_s=5.0):
    """The mosaic / A-projection gridder (mosaic_grid_jit, mosaic_grid.py:
    306-486) on the benchmark's array (64 antennas, 2016 baselines) over
    `n_time` time samples x 64 channels x 4 polarisations, 2048 x 2048
    cube, FP32: 3 baseline groups x 4 channel groups of complex convolution
    functions, 7, 9 or 11 taps per axis, oversampling 8 x 8.

    Algorithmic bytes: visibility 8 + weight 4 bytes per visibility, the
    convolution functions once, the grid written once (8 bytes per cell)."""
    import torch

    from astroviper_amd import _lib, synth
    from astroviper_amd import device as D
    from astroviper_amd.mosaic_grid import _plan_shape

    cfg = synth.CONFIG2
    arr = synth.ArrayConfig(**{**cfg["array"].__dict__, "n_time": n_time})
    sky = cfg["sky"]
    n_uv = cfg["n_uv"]
    cv = cfg["cell_arcsec"] * synth.ARCSEC
    cell = (-cv, cv)
    obs = synth.make_observation(arr, sky)
    uvw, vis, w, flag = synth.device_visibilities(obs, sky, cfg["flag_frac"], cfg["weight_range"],
                                                  device=dev)
    rows, nchan, npol = vis.shape
    nb = obs.uvw.shape[1]
    V = rows * nchan * npol
    rng = np.random.default_rng(12)
    ws, cf, pg, ov = _mosaic_cfs(rng)
    ncb, ncc, ncp, cu_n, cv_n = cf.shape
    cfbm = (np.arange(nb) % ncb).astype(np.int32)
    cfcm = (np.arange(nchan) * ncc // nchan).astype(np.int32)
    cfpm = np.zeros(npol, dtype=np.int32)
    smax = int(ws.max())
    support, tile = _plan_shape(smax)
    w_eff = torch.where(flag.bool(), torch.zeros_like(w), w)  # flagged samples weigh 0
    freq = torch.from_numpy(obs.freq).to(dev)

    def grid_on(rows_, vis_, w_, nch, chan_ids):
        plan = D.GridPlan(rows_, nch, nch, n_uv, cell, support, ov[0], tile=(tile, tile),
                          device=dev)
        plan.set_uvw(uvw[:rows_].contiguous(), freq[chan_ids].contiguous(),
                     torch.arange(nch, dtype=torch.int32, device=dev))
        i32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(dev)
        keep = (i32(cfbm), i32(cfcm[chan_ids.cpu().numpy()]), i32(cfpm), i32(ws))
        cfpg = (torch.from_numpy(cf).to(dev) * torch.from_numpy(pg[0]).to(dev)).to(
            torch.complex64).contiguous()
        g = torch.zeros((nch, npol) + tuple(n_uv), dtype=torch.complex64, device=dev)
        s = torch.zeros((nch, npol), dtype=torch.float64, device=dev)

        def call():
            _lib.call("av_mosaic_grid", plan._h, D._ptr(vis_), D._ptr(w_), npol, nb,
                      D._ptr(keep[0]), D._ptr(keep[1]), D._ptr(keep[2]), D._ptr(cfpg), ncb, ncc, ncp,
                      cu_n, cv_n, D._ptr(keep[3]), ov[0], ov[1], D._ptr(g), D._ptr(s),
                      _lib.AV_F32, 0, D._stream())
        return plan, g, s, call

    allc = torch.arange(nchan, device=dev)
    plan, g, s, call = grid_on(rows, vis.reshape(rows * nchan, npol),
                               w_eff.reshape(rows * nchan, npol), nchan, allc)
    ms = _events_ms(call)
    log(f"bench: mosaic {V} vis in {ms:.2f} ms on the GPU")
    st = plan.stats()
    plan.close()
    cells = nchan * npol * n_uv[0] * n_uv[1]
    nbytes = V * 12 + cf.size * 8 + cells * 8
    out = {"row": "mosaic", "workload": f"mosaic gridder, 64 antennas x {n_time} times x "
                                         f"{nchan} chan x {npol} pol = {V} vis, "
                                         f"{n_uv[0]}x{n_uv[1]} cube, FP32, {ncb}x{ncc} "
                                         f"convolution functions of 7-11 taps per axis, "
                                         f"oversampling {ov[0]}x{ov[1]}, {tile}x{tile} tiles",
           "value": round(V / (ms / 1e3) / 1e6, 1), "unit": "M visibilities/sec",
           "gpu_ms": round(ms, 3), "records": st.get("records"),
           "roofline": _roof(nbytes, ms, "k_mosaic_grid")}
    del g
    # CPU leg: the FP64 oracle on the first rows x 4 channels, one core; the
    # GPU grids the same sample in FP32 for the comparison
    from oracle import oracle as O
    nch_s = 4
    cs = torch.arange(nch_s, device=dev)
    nt_s = 2
    while True:
        r_s = nt_s * nb
        u_h = uvw[:r_s].cpu().numpy().reshape(nt_s, nb, 3)
        v_h = vis[:r_s, :nch_s].cpu().numpy().astype(np.complex128).reshape(nt_s, nb, nch_s, npol)
        w_h = w_eff[:r_s, :nch_s].cpu().numpy().astype(np.float64).reshape(nt_s, nb, nch_s, npol)
        gref = np.zeros((nch_s, npol) + tuple(n_uv), dtype=np.complex128)
        sref = np.zeros((nch_s, npol))
        t0 = time.perf_counter()
        O.mosaic_grid(gref, sref, False, v_h, u_h, obs.freq[:nch_s], np.arange(nch_s),
                      np.arange(npol), cfbm, cfcm[:nch_s], cfpm, w_h, cf, np.array(n_uv), cell,
                      ws, np.array(ov), pg)
        tc = time.perf_counter() - t0
        log(f"bench: mosaic CPU leg {r_s} rows: {tc:.2f} s")
        if tc > cpu_s / 4 or r_s * 4 > rows:
            break
        nt_s *= 4
    vs = vis[:r_s, :nch_s].contiguous().reshape(r_s * nch_s, npol)
    wsm = w_eff[:r_s, :nch_s].contiguous().reshape(r_s * nch_s, npol)
    plan, g, s, call = grid_on(r_s, vs, wsm, nch_s, cs)
    call()
    gg = g.cpu().numpy()
    plan.close()
    out["cpu_baseline"] = {"value": round(r_s * nch_s * npol / tc / 1e6, 3),
                           "unit": "M visibilities/sec", "cores": 1, "kind": "port",
                           "sample": f"first {r_s} rows x {nch_s} channels x {npol} pols, "
                                     f"FP64: {tc:.2f} s"}
    out["parity_vs_cpu"] = {"visibilities": r_s * nch_s * npol,
                            "grid_max_rel": float(np.max(np.abs(gg - gref)) / np.max(np.abs(gref))),
                            "sumwt_max_rel": float(np.max(np.abs(s.cpu().numpy() - sref))
                                                   / np.max(np.abs(sref)))}
    return out


def row_major_cycle(dev, reps=3):
    """One major cycle of imager.py:108-342 on the benchmark workload (64
    antennas x 1000 times x 64 channels x 4 pols, 2048 x 2048 cube, FP32
    gridding and prediction, FP64 FFTs for the dirty image) with every
    visibility-sized array and the binning
    plan resident on the GPU: grid the residual visibilities and form the
    dirty image (grid2image_spheroid_ms4), FFT a model image to the uv plane
    and degrid it (predict_model_visibilities, imager.py:88-105), and form
    RESIDUAL = VISIBILITY - VISIBILITY_MODEL (imager.py:82-85).  CLEAN is
    the Hogbom row.  Beside it: the same cycle when the visibilities,
    weights and flags are staged from pinned host memory and the model and
    residual visibilities copied back each cycle, as a host-resident
    caller must."""
    import torch

    from astroviper_amd import synth
    from astroviper_amd.device import GridPlan, dirty_image
    from astroviper_amd.major_cycle import model_to_uv

    cfg = synth.CONFIG2
    arr = cfg["array"]
    sky = cfg["sky"]
    n_uv = cfg["n_uv"]
    cv = cfg["cell_arcsec"] * synth.ARCSEC
    cell = (-cv, cv)
    obs = synth.make_observation(arr, sky)
    uvw, vis, w, flag = synth.device_visibilities(obs, sky, cfg["flag_frac"], cfg["weight_range"],
                                                  device=dev)
    rows, nchan, npol = vis.shape
    V = rows * nchan * npol
    freq = torch.from_numpy(obs.freq).to(dev)
    vis2, w2, fl2 = (t.reshape(rows * nchan, npol) for t in (vis, w, flag))
    plan = GridPlan(rows, nchan, nchan, n_uv, cell, 7, 100, tile=(16, 16), device=dev,
                    n_baseline=obs.uvw.shape[1])
    plan.set_uvw(uvw, freq)
    grid = torch.empty((nchan, npol) + tuple(n_uv), dtype=torch.complex64, device=dev)
    sumwt = torch.empty((nchan, npol), dtype=torch.float64, device=dev)
    model_vis = torch.zeros_like(vis2)
    resid = vis2.clone()
    # a model image: 0.5 of each plane's two brightest dirty pixels
    plan.grid(vis2, w2, fl2, grid=grid, sumwt=sumwt)
    d0 = dirty_image(grid, sumwt)
    model = torch.zeros_like(d0)
    flat = d0.reshape(nchan * npol, -1)
    top = torch.topk(flat, 2, dim=1)
    model.reshape(nchan * npol, -1).scatter_(1, top.indices, 0.5 * top.values)
    del d0, flat
    t = {}

    def cycle(ev=None):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
        e[0].record()
        plan.grid(resid, w2, fl2, grid=grid, sumwt=sumwt)
        e[1].record()
        img = dirty_image(grid, sumwt)
        e[2].record()
        uvm = model_to_uv(model, torch.complex64)  # as ResidentImager predicts in FP32
        plan.degrid(uvm, fl2, model_vis)
        e[3].record()
        torch.sub(vis2, model_vis, out=resid)
        e[4].record()
        e[4].synchronize()
        for name, a, b in (("grid", 0, 1), ("image", 1, 2), ("predict", 2, 3),
                           ("residual", 3, 4)):
            t.setdefault(name, []).append(e[a].elapsed_time(e[b]))
        del img, uvm

    cycle()
    t.clear()
    for _ in range(reps):
        cycle()
        log(f"bench: major cycle {sum(v[-1] for v in t.values()):.2f} ms")
    parts = {k: round(float(np.median(v)), 3) for k, v in t.items()}
    tot = sum(parts.values())
    # the same cycle with host-resident data: H2D of visibilities, weights
    # and flags, D2H of the model and residual visibilities (pinned buffers)
    h_vis = torch.empty(vis2.shape, dtype=vis2.dtype, pin_memory=True)
    h_w = torch.empty(w2.shape, dtype=w2.dtype, pin_memory=True)
    h_fl = torch.empty(fl2.shape, dtype=fl2.dtype, pin_memory=True)
    h_mod = torch.empty(vis2.shape, dtype=vis2.dtype, pin_memory=True)
    h_res = torch.empty(vis2.shape, dtype=vis2.dtype, pin_memory=True)
    h_vis.copy_(vis2)
    h_w.copy_(w2)
    h_fl.copy_(fl2)

    def transfers():
        vis2.copy_(h_vis, non_blocking=True)
        w2.copy_(h_w, non_blocking=True)
        fl2.copy_(h_fl, non_blocking=True)
        h_mod.copy_(model_vis, non_blocking=True)
        h_res.copy_(resid, non_blocking=True)

    t_x = _events_ms(transfers, reps=reps)
    xbytes = (vis2.numel() * vis2.element_size() * 3 + w2.numel() * w2.element_size()
              + fl2.numel())
    plan.close()
    b_cycle = V * 22 + rows * 48 + nchan * npol * n_uv[0] * n_uv[1] * 16
    return {"row": "major_cycle",
            "workload": f"one major cycle without CLEAN, {V} vis resident (64 antennas x 1000 "
                        f"times x {nchan} chan x {npol} pol), {n_uv[0]}x{n_uv[1]} cube, FP32 "
                        f"gridding and prediction, FP64 FFTs for the dirty image",
            "value": round(V / (tot / 1e3) / 1e6, 1), "unit": "M visibilities/sec",
            "ms": {**parts, "cycle": round(tot, 3)},
            "roofline": _roof(b_cycle, parts["grid"] + parts["predict"], "grid + degrid"),
            "host_resident_cycle": {"transfer_ms": round(t_x, 3),
                                    "transfer_GB": round(xbytes / 1e9, 2),
                                    "host_GB_per_s": round(xbytes / (t_x / 1e3) / 1e9, 2),
                                    "cycle_ms": round(tot + t_x, 3),
                                    "M_vis_per_s": round(V / ((tot + t_x) / 1e3) / 1e6, 1)}}


def row_gcf(dev, n=1024, n_chan=64, cpu_s=5.0):
    """The mosaic gridder's convolution functions
    (make_gridding_convolution_function.py:24-193) for a 2-type array (3
    antenna-type pairs), 64 channels over 1.0-1.4 GHz (57 primary-beam
    channels at the default tolerance), a 1024 x 1024 padded grid,
    oversampling 10: aperture patterns (voltage and power), baseline
    products, FFTs, supports, crops and the phase gradient.

    Algorithmic bytes: each of the two pattern sets is written once by the
    pattern kernel and read and written by each FFT pass (counted as one
    read and one write of the complex plane)."""
    import torch
    from types import SimpleNamespace

    from astroviper_amd.make_gridding_convolution_function import \
        make_gridding_convolution_function

    n_ant = 12
    names = np.array([f"A{i:02d}" for i in range(n_ant)])
    bl = np.array([(i, j) for i in range(n_ant) for j in range(i + 1, n_ant)])
    freq = np.linspace(1.0e9, 1.4e9, n_chan)
    field = SimpleNamespace(FIELD_PHASE_CENTER_DIRECTION=np.array([1.0 + 1e-3, 0.5]))
    ms = SimpleNamespace(xr_ms=SimpleNamespace(get_field_and_source_xds=lambda name: field),
                         antenna_xds=SimpleNamespace(antenna_name=names),
                         baseline_antenna1_name=names[bl[:, 0]],
                         baseline_antenna2_name=names[bl[:, 1]], frequency=freq,
                         polarization=np.array(["XX", "YY"]))
    gcf = {"function": "casa_airy", "list_dish_diameters": np.array([10.7, 6.0]),
           "list_blockage_diameters": np.array([0.75, 0.0]),
           "unique_ant_indx": (np.arange(n_ant) % 3 == 2).astype(int),
           "phase_direction": np.array([1.0, 0.5])}
    gp = {"image_size": np.array([n, n]), "image_size_padded": np.array([n, n]),
          "cell_size": np.array([-3.5e-5, 3.5e-5])}
    run = lambda: make_gridding_convolution_function(None, ms, gcf, gp, {})
    out = run()
    torch.cuda.synchronize()
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        out = run()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    npair, npb = out["CONV_KERNEL"].data.shape[:2]
    planes = npair * npb
    nbytes = 2 * planes * n * n * (8 + 2 * 16)
    res = {"row": "gcf", "workload": f"convolution functions, {npair} antenna-type pairs x {npb} "
                                      f"primary-beam channels, {n}x{n} padded grid, "
                                      f"oversampling 10, casa_airy",
           "value": round(planes / t, 1), "unit": "kernel planes/sec (voltage + power)",
           "ms": round(t * 1e3, 2), "supports": sorted(set(
               int(v) for v in np.asarray(out["SUPPORT"].data).ravel())),
           "roofline": _roof(nbytes, t * 1e3, "patterns + FFT + crop (whole call)")}
    # CPU leg: the oracle's restatement (NumPy/SciPy, one thread for the
    # FFT) on the first antenna-type pair and the first primary-beam
    # channels, and the GPU planes compared with it
    from oracle import oracle as O
    g = dict(gcf, oversampling=np.array([10, 10]), max_support=np.array([15, 15]),
             support_cut_level=2.5e-2)
    g["resize_conv_size"] = (g["max_support"] + 1) * g["oversampling"]
    cmap, pbf = O.gcf_cf_chan_map(freq, 0.005)
    k = 1
    while True:
        t0 = time.perf_counter()
        ks = []
        for ipower in (1, 2):
            g["ipower"] = ipower
            bp = O.gcf_baseline_patterns(pbf[:k], [0], np.array([[0, 0]]), True, g, gp)
            ks.append(np.real(np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(
                bp, axes=(3, 4)), axes=(3, 4)), axes=(3, 4))))
        rk, rkc, sup = O.gcf_resize_and_calc_support(ks[0], ks[1], g, gp)
        tc = time.perf_counter() - t0
        if tc > cpu_s / 4 or 2 * k > len(pbf):
            break
        k *= 2
    got = np.asarray(out["CONV_KERNEL"].data)[0, :k]
    res["cpu_baseline"] = {"value": round(k / tc, 3), "unit": "kernel planes/sec (voltage + power)",
                           "cores": 1, "kind": "port",
                           "sample": f"pair (0, 0), first {k} primary-beam channels: {tc:.2f} s"}
    res["parity_vs_cpu"] = {"planes": k,
                            "support_equal": bool(np.array_equal(
                                np.asarray(out["SUPPORT"].data)[0, :k], sup[0])),
                            "conv_kernel_max_rel": float(np.max(np.abs(got - rk[0]))
                                                         / np.max(np.abs(rk[0])))}
    return res


def run_rows(names, dev):
    rows_fn = {"hogbom": lambda: row_hogbom(dev),
               "mosaic": lambda: row_mosaic(dev),
               "gcf": lambda: row_gcf(dev),
               "major_cycle": lambda: row_major_cycle(dev)}
    for name in names:
        if name == "weights":
            import torch

            from astroviper_amd import synth
            cfg = synth.CONFIG2
            obs = synth.make_observation(cfg["array"], cfg["sky"])
            uvw, vis, w, flag = synth.device_visibilities(obs, cfg["sky"], cfg["flag_frac"],
                                                          cfg["weight_range"], device=dev)
            del vis
            cv = cfg["cell_arcsec"] * synth.ARCSEC
            res = row_weights(dev, obs, uvw, w, flag, cfg["n_uv"], (-cv, cv))
            del uvw, w, flag
            torch.cuda.empty_cache()
        elif name in rows_fn:
            res = rows_fn[name]()
        else:
            raise SystemExit(f"unknown row {name!r}")
        log(f"bench: row {name} done")
        print(json.dumps(res), flush=True)
        import torch
        torch.cuda.empty_cache()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", choices=["cube", "continuum"], default="cube")
    ap.add_argument("--tile", type=int, nargs=2, default=[16, 16])
    ap.add_argument("--unit-cap", type=int, default=0)
    ap.add_argument("--compensated", action="store_true")
    ap.add_argument("--subtile", action="store_true",
                    help="grid FP32 16x16 tiles with the sub-tile scheduled 4x4x1 kernel")
    ap.add_argument("--f32-products", action="store_true",
                    help="grid and degrid FP32 16x16 tiles with FP32 products (16x16x4 "
                         "kernels) instead of split-FP16 operands")
    ap.add_argument("--row-order-off", action="store_true",
                    help="bin the rows in row order instead of baseline by baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--time-samples", type=int, default=1000)
    ap.add_argument("--rows", default="",
                    help="comma list of rows beside the hot path to measure instead of the "
                         "headline: hogbom, weights, mosaic, gcf, major_cycle")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="after the timed steps, copy a BYTES-sized device buffer once "
                         "(the reference kernel for scaling HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import torch

    if args.rows:
        run_rows([r for r in args.rows.split(",") if r], torch.device("cuda", 0))
        return

    from astroviper_amd import synth
    from astroviper_amd.device import (GridPlan, degrid_kernel_name, degrid_plain,
                                       grid_kernel_name, grid_plain)

    dev = torch.device("cuda", 0)
    cfg = synth.CONFIG2
    arr = synth.ArrayConfig(**{**cfg["array"].__dict__, "n_time": args.time_samples})
    sky = cfg["sky"]
    n_uv = cfg["n_uv"]
    cellv = cfg["cell_arcsec"] * synth.ARCSEC
    cell = (-cellv, cellv)
    cube = args.mode == "cube"
    log(f"bench: generating workload ({arr.n_ant} antennas, {arr.n_time} times, "
        f"{arr.n_chan} channels, {arr.n_pol} pols) on the GPU")
    obs = synth.make_observation(arr, sky)
    uvw, vis, w, flag = synth.device_visibilities(obs, sky, cfg["flag_frac"],
                                                  cfg["weight_range"], device=dev)
    rows, nchan, npol = vis.shape
    V = rows * nchan * npol
    n_img = nchan if cube else 1
    freq = torch.from_numpy(obs.freq).to(dev)
    vis2 = vis.reshape(rows * nchan, npol)
    w2 = w.reshape(rows * nchan, npol)
    fl2 = flag.reshape(rows * nchan, npol)
    plan = GridPlan(rows, nchan, n_img, n_uv, cell, 7, 100, tile=tuple(args.tile),
                    unit_cap=args.unit_cap, device=dev,
                    n_baseline=1 if args.row_order_off else obs.uvw.shape[1])
    grid = torch.empty((n_img, npol) + tuple(n_uv), dtype=torch.complex64, device=dev)
    sumwt = torch.empty((n_img, npol), dtype=torch.float64, device=dev)
    model = torch.empty((rows * nchan, npol), dtype=torch.complex64, device=dev)
    torch.cuda.synchronize()
    log(f"bench: {V} visibilities resident; {n_img * npol} planes of {n_uv[0]}x{n_uv[1]}")

    def step(ev=None):
        if ev is not None:
            ev[0].record()
        plan.set_uvw(uvw, freq)
        if ev is not None:
            ev[1].record()
        plan.grid(vis2, w2, fl2, grid=grid, sumwt=sumwt, compensated=args.compensated,
                  subtile=args.subtile, f32_products=args.f32_products)
        if ev is not None:
            ev[2].record()
        plan.degrid(grid, fl2, model, f32_products=args.f32_products)
        if ev is not None:
            ev[3].record()

    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"bench: warmup step {i + 1}/{args.warmup} done")
    st = plan.stats()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    t_plan = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    t_grid = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
    t_degrid = float(np.mean([e[2].elapsed_time(e[3]) for e in evs]))
    value = V / (ms / 1e3) / 1e6
    log(f"bench: {args.steps} steps, {ms:.2f} ms/step (plan {t_plan:.2f}, grid {t_grid:.2f}, "
        f"degrid {t_degrid:.2f} ms) -> {value:.0f} M vis/s")

    cells = n_img * npol * n_uv[0] * n_uv[1]
    b_grid = V * 13 + cells * 8          # vis 8 + weight 4 + flag 1; grid written once
    b_degrid = V * 9 + cells * 8         # flag 1 + model 8; grid read once
    b_plan = rows * 24 + nchan * 8       # uvw once, frequencies
    b_pair = V * 22 + rows * 48 + cells * 16 + nchan * 16
    kern = {"grid": (t_grid, b_grid), "degrid": (t_degrid, b_degrid)}
    dom = max(kern, key=lambda k: kern[k][0])
    tdom, bdom = kern[dom]
    achieved = bdom / (tdom / 1e3) / 1e9
    names = {"grid": grid_kernel_name(args.tile, npol, compensated=args.compensated,
                                      subtile=args.subtile, f32_products=args.f32_products),
             "degrid": degrid_kernel_name(args.tile, npol, 7, f32_products=args.f32_products)}
    roofline = {"bound": "hbm", "kernel": names[dom],
                "achieved": round(achieved, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": round(achieved * 1e9 / HBM_PEAK, 4),
                "traffic": measured_traffic(names[dom], V, n_uv, args.mode, args.tile),
                "algorithmic_bytes": bdom, "avg_ms": round(tdom, 3)}

    extra = {"plan_ms": round(t_plan, 3), "grid_ms": round(t_grid, 3),
             "degrid_ms": round(t_degrid, 3),
             "pair_roofline_frac": round(b_pair / (ms / 1e3) / HBM_PEAK, 4),
             "pair_algorithmic_bytes": b_pair, "plan_stats": st}

    # plain one-thread-per-sample gridder on a channel subset, scaled
    if not args.no_plain:
        chans = list(range(0, nchan, max(1, nchan // 8)))
        cmap = list(range(nchan)) if cube else [0] * nchan
        pg = torch.zeros_like(grid)
        psw = torch.zeros_like(sumwt)
        pm = torch.zeros((rows, nchan, npol), dtype=torch.complex64, device=dev)
        cgk = plan.cgk
        grid_plain(uvw, freq, vis, w, flag, n_uv, cell, n_img, cmap, grid=pg, sumwt=psw,
                   cgk=cgk, chans=chans[:1])
        torch.cuda.synchronize()
        ta = time.perf_counter()
        pg.zero_()
        grid_plain(uvw, freq, vis, w, flag, n_uv, cell, n_img, cmap, grid=pg, sumwt=psw,
                   cgk=cgk, chans=chans)
        degrid_plain(uvw, freq, pg, flag, cell, cmap, npol, vis_out=pm, cgk=cgk, chans=chans)
        torch.cuda.synchronize()
        tb = time.perf_counter()
        vsub = rows * len(chans) * npol
        plain = vsub / (tb - ta) / 1e6
        extra["plain_gpu"] = {"value": round(plain, 1), "unit": "M visibilities/sec",
                              "sample": f"{len(chans)} of {nchan} channels, one launch each"}
        extra["speedup_vs_plain_gpu"] = round(value / plain, 2)
        log(f"bench: plain GPU gridder {plain:.0f} M vis/s on {len(chans)} channels")
        del pg, pm

    # accuracy of this FP32 run against an FP64 GPU run on the same inputs
    if not args.no_check:
        g64, sw64 = plan.grid(vis2.to(torch.complex128), w2.to(torch.float64), fl2)
        m64 = plan.degrid(g64, fl2)
        torch.cuda.synchronize()

        def relerr(a, b):
            return float((a - b).abs().max() / b.abs().max())

        errs = {
            "grid": relerr(grid.to(torch.complex128), g64),
            "sumwt": relerr(sumwt, sw64),
            "degrid_vs_fp64_grid": relerr(model.to(torch.complex128), m64),
        }
        del m64
        # dirty images and PSFs of the first image channel's planes (FFT in
        # FP64 for both, as grid2image_spheroid_ms4 does)
        from astroviper_amd.device import dirty_image, ps_correction_1d
        tu = ps_correction_1d(n_uv[0], dev)
        tv = ps_correction_1d(n_uv[1], dev)
        inner = (tu[:, None] * tv[None, :]) >= 1e-2

        def img_err(a, b):
            d = (a - b).abs()
            m = b.abs().max()
            return {"full": float(d.max() / m), "taper_ge_1e-2": float(d[..., inner].max() / m)}

        errs["dirty"] = img_err(dirty_image(grid[:1], sumwt[:1]), dirty_image(g64[:1], sw64[:1]))
        del g64
        p32, ps32 = plan.grid(None, w2, fl2, psf=True, compensated=args.compensated,
                              subtile=args.subtile, f32_products=args.f32_products)
        p64, ps64 = plan.grid(None, w2.to(torch.float64), fl2, psf=True)
        errs["psf"] = img_err(dirty_image(p32[:1], ps32[:1]), dirty_image(p64[:1], ps64[:1]))
        errs["images_of"] = f"image channel 0 ({npol} planes)"
        extra["max_rel_err_vs_fp64"] = errs
        log(f"bench: errors vs FP64 {extra['max_rel_err_vs_fp64']}")
        del p32, p64

    if args.calibrate_copy > 0:
        src = torch.empty(args.calibrate_copy // 4, dtype=torch.float32, device=dev)
        src.fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(obs, uvw, vis, w, flag, n_uv, cell, cube)
        log(f"bench: cpu baseline {cpu['value']:.2f} M vis/s ({cpu['sample']}); one core "
            f"{cpu['single_core']['value']:.2f} M vis/s")
        extra["small_workload"] = small_workload_leg(dev)
        log(f"bench: 27-antenna workload {extra['small_workload']}")

    out = {
        "metric": "M visibilities/sec (grid+degrid, support 7) batched FP32 vs plain GPU @<=1e-4 rel err",
        "value": round(value, 1),
        "unit": "M visibilities/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 coordinates and weight sums)",
        "data": "synthetic (seeded point-source sky, earth-rotation uvw)",
        "config": {"workload": f"standard gridder grid+degrid, {arr.n_ant} antennas x "
                               f"{arr.n_time} times x {nchan} chan x {npol} pol = {V} vis, "
                               f"{n_uv[0]}x{n_uv[1]} {args.mode}, support 7, oversampling 100, "
                               f"10% flagged",
                   "mode": args.mode, "tile": list(args.tile), "visibilities": V,
                   "compensated": bool(args.compensated), "subtile": bool(args.subtile),
                   "row_order": "row" if args.row_order_off else "baseline",
                   "grid_kernel": names["grid"], "degrid_kernel": names["degrid"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
