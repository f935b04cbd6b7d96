"""Precision sweep of the batched gridder on one GPU.

Grids the same device-resident visibilities three ways (FP32 accumulation,
FP32 with compensated per-tile sums, FP64) and reports, against the FP64
result, the max and RMS relative error of the grid, the dirty image and the
PSF, the relative error of the sums of weights, and visibilities/s of each
variant's grid+degrid pair.  Coordinates and sums of weights are FP64 in
every variant.  Dirty images follow grid2image_spheroid_ms4
(core/imaging/imaging_utils/standard_grid.py:693-708) and are formed a few
planes at a time so that the FP64 FFT work buffer stays small.

    python -m astroviper_amd.sweep [--n 4096] [--n-time 1938] [--chan 64]

The default is 64 antennas x 1938 times x 64 channels x 4 polarisations =
1,000,144,896 visibilities on a 4096 x 4096 cube grid with 1" cells.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import torch

from . import synth
from .device import GridPlan, dirty_image, require_gpu


def _errs(x: torch.Tensor, ref: torch.Tensor, planes_dim: int = 2, chunk: int = 4):
    """(max over planes of max|x-ref|/max|ref|, global RMS relative), over
    `chunk` entries of the leading axis at a time."""
    worst, num, den = 0.0, 0.0, 0.0
    for c0 in range(0, x.shape[0], chunk):
        xs, rs = x[c0:c0 + chunk], ref[c0:c0 + chunk]
        d = (xs.to(rs.dtype) - rs).abs().flatten(planes_dim)
        r = rs.abs().flatten(planes_dim)
        rmax = r.amax(dim=-1)
        per = d.amax(dim=-1) / torch.where(rmax > 0, rmax, torch.ones_like(rmax))
        worst = max(worst, float(per.max()))
        num += float((d.double() ** 2).sum())
        den += float((r.double() ** 2).sum())
        del d, r
    return worst, math.sqrt(num / max(den, 1e-300))


def _image_errs(g: torch.Tensor, sw: torch.Tensor, gref: torch.Tensor, swref: torch.Tensor,
                chunk: int):
    """Dirty-image errors over all planes, `chunk` image channels at a time,
    plus the errors where the taper correction is at most 100x."""
    from .device import ps_correction_1d
    n_u, n_v = g.shape[-2:]
    tu = ps_correction_1d(n_u, g.device)
    tv = ps_correction_1d(n_v, g.device)
    inner = (tu[:, None] * tv[None, :]) >= 1e-2
    worst, worst_in, num, den = 0.0, 0.0, 0.0, 0.0
    for c0 in range(0, g.shape[0], chunk):
        c1 = min(g.shape[0], c0 + chunk)
        img = dirty_image(g[c0:c1], sw[c0:c1])
        ref = dirty_image(gref[c0:c1], swref[c0:c1])
        d = (img - ref).abs()
        m = ref.abs().flatten(2).amax(dim=-1)
        m = torch.where(m > 0, m, torch.ones_like(m))
        worst = max(worst, float((d.flatten(2).amax(dim=-1) / m).max()))
        worst_in = max(worst_in, float((d[..., inner].amax(dim=-1) / m).max()))
        num += float((d ** 2).sum())
        den += float((ref ** 2).sum())
        del img, ref, d
    return worst, math.sqrt(num / max(den, 1e-300)), worst_in


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--n-time", type=int, default=1938)
    ap.add_argument("--chan", type=int, default=64)
    ap.add_argument("--ant", type=int, default=64)
    ap.add_argument("--cell-arcsec", type=float, default=1.0)
    ap.add_argument("--tile", type=int, nargs=2, default=[16, 16])
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--image-chunk", type=int, default=4)
    args = ap.parse_args(argv)
    dev = require_gpu()

    def log(msg):
        print(f"sweep: {msg}", file=sys.stderr, flush=True)

    arr = synth.ArrayConfig(n_ant=args.ant, n_time=args.n_time, n_chan=args.chan, n_pol=4, seed=2)
    sky = synth.SkyConfig()
    obs = synth.make_observation(arr, sky)
    uvw, vis, w, flag = synth.device_visibilities(obs, sky, 0.1, (0.5, 1.5), device=dev)
    rows, nchan, npol = vis.shape
    V = rows * nchan * npol
    log(f"{V} visibilities resident; grid {nchan}x{npol}x{args.n}^2")
    freq = torch.from_numpy(obs.freq).to(dev)
    c = args.cell_arcsec * synth.ARCSEC
    n_uv = (args.n, args.n)
    plan = GridPlan(rows, nchan, nchan, n_uv, (-c, c), 7, 100, tile=tuple(args.tile), device=dev,
                    n_baseline=obs.uvw.shape[1])
    plan.set_uvw(uvw, freq)
    v2, w2, f2 = vis.reshape(-1, npol), w.reshape(-1, npol), flag.reshape(-1, npol)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.reps

    # FP64 reference on the same (single-precision) samples
    g64 = torch.empty((nchan, npol) + n_uv, dtype=torch.complex128, device=dev)
    sw64 = torch.empty((nchan, npol), dtype=torch.float64, device=dev)
    v64, w64 = v2.to(torch.complex128), w2.to(torch.float64)
    out64 = torch.empty_like(v64)
    t64 = timed(lambda: (plan.grid(v64, w64, f2, grid=g64, sumwt=sw64),
                         plan.degrid(g64, f2, out64)))
    log(f"fp64 pair {t64:.1f} ms")
    res = {"fp64": {"ms": round(t64, 2), "M_vis_per_s": round(V / t64 / 1e3, 1)}}
    del v64, w64

    g32 = torch.empty((nchan, npol) + n_uv, dtype=torch.complex64, device=dev)
    sw32 = torch.empty((nchan, npol), dtype=torch.float64, device=dev)
    out32 = torch.empty_like(v2)
    for name, comp in (("fp32", False), ("fp32_compensated", True)):
        t = timed(lambda: (plan.grid(v2, w2, f2, grid=g32, sumwt=sw32, compensated=comp),
                           plan.degrid(g32, f2, out32)))
        ge = _errs(g32, g64)
        swe = float(((sw32 - sw64).abs() / sw64.abs().clamp_min(1e-300)).max())
        de = _errs(out32.reshape(1, -1), out64.reshape(1, -1), planes_dim=1)
        ie = _image_errs(g32, sw32, g64, sw64, args.image_chunk)
        res[name] = {"ms": round(t, 2), "M_vis_per_s": round(V / t / 1e3, 1),
                     "grid_max_rel": ge[0], "grid_rms_rel": ge[1],
                     "dirty_max_rel": ie[0], "dirty_rms_rel": ie[1],
                     "dirty_max_rel_taper_ge_1e-2": ie[2],
                     "sumwt_rel": swe, "degrid_max_rel": de[0], "degrid_rms_rel": de[1]}
        log(f"{name}: {res[name]}")
    # PSF: weights only
    plan.grid(None, w2.to(torch.float64), f2, grid=g64, sumwt=sw64, psf=True, dtype=torch.float64)
    for name, comp in (("fp32", False), ("fp32_compensated", True)):
        plan.grid(None, w2, f2, grid=g32, sumwt=sw32, psf=True, compensated=comp)
        pe = _image_errs(g32, sw32, g64, sw64, args.image_chunk)
        res[name].update({"psf_max_rel": pe[0], "psf_rms_rel": pe[1],
                          "psf_max_rel_taper_ge_1e-2": pe[2]})
    print(json.dumps({"workload": f"{args.ant} antennas x {args.n_time} times x {nchan} chan x "
                                  f"{npol} pol = {V} vis, {args.n}x{args.n} cube, "
                                  f"{args.cell_arcsec}\" cells, support 7, 10% flagged",
                      "visibilities": V, "tile": args.tile, "variants": res}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
