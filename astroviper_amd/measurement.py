"""Differentiable measurement operator: sky image -> visibilities.

    A x = D F (x / t)

t is the gridding-correction taper (the outer product of grdsf on both
axes, gcf_prolate_spheroidal.py:72-79), F = fftshift . fft2 . ifftshift
(fft_lm_to_uv, core/imaging/fft.py:5-18) and D the normalised degridder
(dgrid_numba, standard_degrid.py:388-485).  The reference's predictor
(imager.py:88-105) is F then D without the taper; dividing by the taper
first makes D F (x / t) the model visibilities of the sky x.

The adjoint is A^H y = (F^H D^H y) / t with F^H = fftshift . (N^2 ifft2)
. ifftshift and D^H the gridding kernel weighting each sample by 1/norm
(av_grid with AV_GRID_ADJOINT on a plan whose gridding geometry is the
degridder's), so grid and degrid are exact transposes of each other.
In FP32 the grid and degrid kernels run in FP32 (the adjoint with
compensated sums) while the taper and FFT run in FP64: the taper division
reaches 1/grdsf(1)^2 = 6e4 at the image corners and would otherwise
magnify FP32 FFT rounding.
`MeasurementFunction` wraps A as a torch.autograd.Function whose backward is
that adjoint (the real part for a real image).
"""
from __future__ import annotations

from typing import Optional

import torch

from . import device as _dev
from ._lib import call


class MeasurementOperator:
    """A = D F (. / t) for fixed uvw; images [n_img_chan, n_pol, n_u, n_v]."""

    def __init__(self, uvw: torch.Tensor, freq: torch.Tensor, n_uv, cell, n_pol: int = 1,
                 chan_map: Optional[torch.Tensor] = None, n_img_chan: int = 1,
                 support: int = 7, oversampling: int = 100, dtype=torch.complex64,
                 tile=(16, 16), device=None):
        self.device = _dev.require_gpu(device)
        self.n_u, self.n_v = int(n_uv[0]), int(n_uv[1])
        self.n_pol = int(n_pol)
        self.n_img_chan = int(n_img_chan)
        self.cdtype = dtype
        self.code = _dev._real_code(dtype)
        self.fft_dtype = torch.complex128
        self.fcode = _dev._real_code(self.fft_dtype)
        self.compensated = self.code == _dev._real_code(torch.complex64)
        rows = uvw.shape[0]
        nchan = freq.numel()
        self.plan = _dev.GridPlan(rows, nchan, n_img_chan, n_uv, cell, support, oversampling,
                                  tile=tile, geometry="degrid", device=self.device)
        self.plan.set_uvw(uvw, freq, chan_map)
        self.tu = _dev.ps_correction_1d(self.n_u, self.device)
        self.tv = _dev.ps_correction_1d(self.n_v, self.device)
        self.shape = (self.n_img_chan, self.n_pol, self.n_u, self.n_v)
        self.n_samples = self.plan.n_samples

    def _pass(self, src, dst, shift_u, shift_v, taper, scale, src_real, dst_real,
              src_code, dst_code):
        planes = src.numel() // (self.n_u * self.n_v)
        call("av_shift_taper", _dev._ptr(src), _dev._ptr(dst), planes, self.n_u, self.n_v,
             shift_u, shift_v, _dev._ptr(self.tu), _dev._ptr(self.tv), taper, float(scale),
             src_code, int(src_real), dst_code, int(dst_real), _dev._stream())

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """Model visibilities [samples, n_pol] of image x (real or complex)."""
        x = x.reshape(self.shape).contiguous()
        real = not x.is_complex()
        rdt = _dev.real_dtype(self.code)
        x = x.to(rdt if real else self.cdtype)
        w = torch.empty(self.shape, dtype=self.fft_dtype, device=self.device)
        # ifftshift, dividing by the taper at the source pixel
        self._pass(x, w, self.n_u // 2, self.n_v // 2, 1, 1.0, real, False, self.code,
                   self.fcode)
        torch.fft.fft2(w, dim=(-2, -1), out=w)
        g = torch.empty(self.shape, dtype=self.cdtype, device=self.device)
        self._pass(w, g, self.n_u - self.n_u // 2, self.n_v - self.n_v // 2, 0, 1.0, False, False,
                   self.fcode, self.code)
        return self.plan.degrid(g, n_pol=self.n_pol)

    def adjoint(self, y: torch.Tensor, real: bool = False) -> torch.Tensor:
        """A^H y for visibilities y [samples, n_pol]; real part if `real`."""
        y = y.reshape(self.n_samples, self.n_pol).to(self.cdtype).contiguous()
        g, _ = self.plan.grid(y, None, adjoint=True, n_pol=self.n_pol,
                              compensated=self.compensated)
        w = torch.empty(self.shape, dtype=self.fft_dtype, device=self.device)
        self._pass(g, w, self.n_u // 2, self.n_v // 2, 0, 1.0, False, False, self.code,
                   self.fcode)
        torch.fft.ifft2(w, dim=(-2, -1), norm="forward", out=w)  # N^2 * ifft2
        out = torch.empty(self.shape, dtype=_dev.real_dtype(self.code) if real else self.cdtype,
                          device=self.device)
        self._pass(w, out, self.n_u - self.n_u // 2, self.n_v - self.n_v // 2, 2, 1.0, False,
                   real, self.fcode, self.code)
        return out

    def apply(self, x: torch.Tensor) -> torch.Tensor:
        """Differentiable A x."""
        return MeasurementFunction.apply(x, self)


class MeasurementFunction(torch.autograd.Function):
    """y = A x with the exact adjoint as backward."""

    @staticmethod
    def forward(ctx, x, op: MeasurementOperator):
        ctx.op = op
        ctx.real = not x.is_complex()
        ctx.shape = x.shape
        ctx.dtype = x.dtype
        return op.forward(x.detach())

    @staticmethod
    def backward(ctx, grad_y):
        g = ctx.op.adjoint(grad_y, real=ctx.real)
        return g.reshape(ctx.shape).to(ctx.dtype), None


def _main(argv=None) -> int:
    """Time the measurement operator on a sky of n x n pixels and about
    `points` uvw samples from the 27-antenna, 4-channel generator (hour
    angles extended until the sample count is reached)."""
    import argparse
    import json
    import math

    import numpy as np

    from . import synth

    ap = argparse.ArgumentParser(description=_main.__doc__)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--points", type=int, default=10_000_000)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--dtype", choices=["fp32", "fp64"], default="fp32")
    args = ap.parse_args(argv)
    dev = _dev.require_gpu()
    uvw, freq, cell = synth.measurement_samples(args.points, args.n)
    dtype = torch.complex64 if args.dtype == "fp32" else torch.complex128
    rows = uvw.shape[0]
    op = MeasurementOperator(torch.from_numpy(uvw).to(dev), torch.from_numpy(freq).to(dev),
                             (args.n, args.n), cell, n_pol=1, dtype=dtype, device=dev)
    sky = torch.from_numpy(synth.blob_sky(args.n, 50, np.random.default_rng(3))).to(dev)
    x = sky.to(torch.float32 if args.dtype == "fp32" else torch.float64).requires_grad_(True)
    y0 = op.forward(x.detach())

    def fwd():
        return op.forward(x.detach())

    def grad():
        x.grad = None
        chi2 = (op.apply(x) - y0).abs().pow(2).sum()
        chi2.backward()

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

    tf = timed(fwd)
    tg = timed(grad)
    n_s = op.n_samples
    print(json.dumps({"workload": f"{args.n}x{args.n} sky, {n_s} samples ({rows} rows x "
                                  f"{freq.size} channels), support 7, {args.dtype}",
                      "forward_ms": round(tf, 3), "forward_plus_backward_ms": round(tg, 3),
                      "M_samples_per_s_forward": round(n_s / tf / 1e3, 1),
                      "M_samples_per_s_forward_backward": round(n_s / tg / 1e3, 1)}),
          flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
