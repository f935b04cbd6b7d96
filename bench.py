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

    plan = GridPlan(rows, nchan, nchan, n_uv, cell, 7, 100, tile=(16, 16), device=dev)
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
    return {"workload": f"27 antennas x {nt} times x {nchan} chan x {npol} pol = {V} vis, "
                        f"{n_uv[0]}x{n_uv[1]} cube, 5\" cells, 10% flagged, weights U(0.5,1.5)",
            "cpu_fp64_one_core": {"value": round(V / float(np.median(tcpu)) / 1e6, 3),
                                  "unit": "M visibilities/sec", "median_s": float(np.median(tcpu))},
            "gpu_fp32": {"value": round(V / (tgpu / 1e3) / 1e6, 1), "unit": "M visibilities/sec",
                         "ms": round(tgpu, 4), "what": "plan + grid + degrid, median of events"},
            "max_rel_err_vs_fp64_oracle": {
                "grid": rel(g.cpu().numpy(), gref), "sumwt": rel(sw.cpu().numpy(), swref),
                "dirty": img_err(img, iref), "psf": img_err(psf, psfref),
                "degrid_from_own_grid": rel(mm, mref)}}


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
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--time-samples", type=int, default=1000)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="after the timed steps, copy a BYTES-sized device buffer once "
                         "(the reference kernel for scaling HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import torch

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
                    unit_cap=args.unit_cap, device=dev)
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
                   "grid_kernel": names["grid"], "degrid_kernel": names["degrid"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
