"""Major cycles with the data resident on the GPU.

The reference's imaging loop (core/imaging/imager.py:108-342) calls its
gridder and degridder through host NumPy arrays every major cycle:
grid_visibilities (imager.py:54-79) grids VISIBILITY or RESIDUAL into a
dirty image, predict_model_visibilities (:88-105) FFTs the model image
(no taper) and degrids it into VISIBILITY_MODEL, and
compute_residual_visibilities (:82-85) forms RESIDUAL = VISIBILITY -
VISIBILITY_MODEL; Hogbom CLEAN (deconvolvers/hogbom) deconvolves each plane
in between.  ResidentImager keeps every visibility-sized array -- the
visibilities, weights, flags, model and residual visibilities -- and one
binning plan of the fixed uvw on the GPU for the whole loop, so a cycle moves
no visibility data between host and device.

Semantics follow the reference calls: weights are zero where FLAG is set
or UVW is NaN (grid2image_spheroid_ms4, standard_grid.py:664-668); such rows
are flagged for the degridder (degrid_spheroid_ms4, standard_degrid.py:
557-566); image channels follow chan_mode (standard_grid.py:119-127), and
the degridder's channel map floor(i * nchan_g / nfreq) (standard_degrid.py:
580) equals it for cube and continuum grids.  CLEAN runs per (channel,
polarisation) plane with the channel's first-polarisation PSF, as
deconvolution.py:224-287 does, in float32 by default: the reference's
hogbom_clean passes a float32 empty mask (deconvolution.py:399-412), which
selects the extension's float32 overload for float64 images.
"""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np
import torch

from . import device as D
from .deconvolvers import hogbom as H

_SIGNS = {}


def model_to_uv(model: torch.Tensor, cdt: torch.dtype) -> torch.Tensor:
    """fftshift(fft2(ifftshift(model))) over the last two axes, in the
    complex dtype `cdt`: fft_lm_to_uv as predict_model_visibilities applies
    it (imager.py:96-99, fft.py:5-18).

    When both axes are multiples of 4 the two shifts by N/2 are sign flips:
    with S[i, j] = (-1)^(i + j), fftshift(fft2(ifftshift(x))) = S * fft2(S * x),
    exactly in exact arithmetic (the shift theorem; (-1)^(N0/2 + N1/2) = 1).
    So the image is cast once, flipped in place, transformed and flipped in
    place: two passes of the cube instead of the four of two rolls.

    A real image (the model) goes through a real-to-complex FFT, which
    computes half the spectrum for half the work: the flipped image is
    formed in the real dtype in one pass, transformed, and
    av_hermitian_expand writes the whole spectrum from the half
    (F[-k] = conj(F[k])) with the output flip, in one pass."""
    ny, nx = model.shape[-2:]
    ax = (-2, -1)
    if ny % 4 or nx % 4:
        m = model.to(cdt)
        return torch.fft.fftshift(torch.fft.fft2(torch.fft.ifftshift(m, dim=ax), dim=ax), dim=ax)
    key = (ny, nx, model.device, cdt)
    s = _SIGNS.get(key)
    if s is None:
        rdt = torch.float32 if cdt == torch.complex64 else torch.float64
        i = torch.arange(ny, device=model.device).reshape(-1, 1)
        j = torch.arange(nx, device=model.device).reshape(1, -1)
        s = (1 - 2 * ((i + j) % 2)).to(rdt)
        _SIGNS[key] = s
    if not model.is_complex():
        rdt = s.dtype
        x = torch.empty(model.shape, dtype=rdt, device=model.device)
        torch.mul(model, s, out=x)
        h = torch.fft.rfft2(x, dim=ax)
        del x
        y = torch.empty(model.shape, dtype=cdt, device=model.device)
        D.call("av_hermitian_expand", D._ptr(h), D._ptr(y), model.numel() // (ny * nx), ny, nx,
               1, D._real_code(cdt), D._stream())
        return y
    x = model.to(cdt)
    x.mul_(s)
    y = torch.fft.fft2(x, dim=ax)
    y.mul_(s)
    return y


class ResidentImager:
    """Device-resident state of one measurement set for major cycles.

    precision "fp64" (the reference's: complex128 grids, FP64 gridding and
    degridding) or "fp32" (complex64 visibilities, FP32 accumulation, FP64
    coordinates, sums of weights and FFTs)."""

    def __init__(self, ms4, image_size: Tuple[int, int], cell_size, support: int = 7,
                 oversampling: int = 100, chan_mode: str = "cube", precision: str = "fp64",
                 tile=(16, 16), device=None):
        self.dev = D.require_gpu(device)
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.cdt = torch.complex128 if precision == "fp64" else torch.complex64
        self.rdt = torch.float64 if precision == "fp64" else torch.float32
        self.ny, self.nx = int(image_size[0]), int(image_size[1])
        self.cell = (float(cell_size[0]), float(cell_size[1]))
        self.support, self.oversampling = int(support), int(oversampling)
        uvw = np.asarray(ms4.UVW.data, dtype=np.float64)
        w = np.asarray(ms4.WEIGHT.data)
        flag = np.asarray(ms4.FLAG.data, dtype=bool)
        vis = np.asarray(ms4.VISIBILITY.data)
        nt, nb, nf, npol = w.shape
        self.shape = (nt, nb, nf, npol)
        rows = nt * nb
        nan_rows = np.isnan(uvw[:, :, 0]) | np.isnan(uvw[:, :, 1])
        flag_all = flag | nan_rows[:, :, None, None]
        self.n_img_chan = nf if chan_mode == "cube" else 1
        cmap = np.arange(nf) if chan_mode == "cube" else np.zeros(nf, dtype=int)
        self.freq = np.asarray(ms4.coords["frequency"].values, dtype=np.float64)
        # image arrays are [.., image_size[0], image_size[1]] (imager.py:66-69)
        self.plan = D.GridPlan(rows, nf, self.n_img_chan, (self.ny, self.nx), self.cell,
                               self.support, self.oversampling, tile=tile, device=self.dev,
                               n_baseline=nb)
        self.plan.set_uvw(torch.from_numpy(np.ascontiguousarray(uvw.reshape(rows, 3))),
                          torch.from_numpy(self.freq),
                          torch.from_numpy(cmap.astype(np.int32)))
        ns = rows * nf
        self.vis = torch.from_numpy(np.ascontiguousarray(vis.reshape(ns, npol))).to(
            self.dev, self.cdt)
        self.weight = torch.from_numpy(np.ascontiguousarray(w.reshape(ns, npol))).to(
            self.dev, self.rdt)
        self.flag = torch.from_numpy(np.ascontiguousarray(flag_all.reshape(ns, npol)).astype(
            np.uint8)).to(self.dev)
        self.model_vis = torch.zeros_like(self.vis)
        self.residual_vis = self.vis.clone()
        self._work = None

    # ------------------------------------------------ imager.py mirrors --
    def grid_visibilities(self, column: str = "VISIBILITY") -> torch.Tensor:
        """imager.grid_visibilities (imager.py:54-79) on the device:
        [n_img_chan, n_pol, ny, nx] float64 dirty image of VISIBILITY or
        RESIDUAL."""
        data = {"VISIBILITY": self.vis, "RESIDUAL": self.residual_vis}[column]
        g, sw = self.plan.grid(data, self.weight, self.flag)
        return D.dirty_image(g, sw, self.support, self.oversampling)

    def make_psf(self) -> torch.Tensor:
        """The PSF of make_psf / grid2image_spheroid_ms4(dopsf=True)."""
        g, sw = self.plan.grid(None, self.weight, self.flag, psf=True)
        return D.dirty_image(g, sw, self.support, self.oversampling)

    def predict_model_visibilities(self, model_image, incremental: bool = False) -> None:
        """imager.predict_model_visibilities (imager.py:88-105): FFT the model
        image [n_img_chan, n_pol, ny, nx] to the uv plane (no taper) and
        degrid it into the model visibilities (added to them when
        incremental)."""
        m = torch.as_tensor(model_image).to(self.dev)
        grid = model_to_uv(m, self.cdt)
        self.plan.degrid(grid.contiguous(), self.flag, self.model_vis, incremental=incremental)

    def compute_residual_visibilities(self) -> None:
        """imager.compute_residual_visibilities (imager.py:82-85):
        RESIDUAL = VISIBILITY - VISIBILITY_MODEL."""
        torch.sub(self.vis, self.model_vis, out=self.residual_vis)

    # ----------------------------------------------------- deconvolution --
    def deconvolve(self, dirty: torch.Tensor, psf: torch.Tensor, niter: int, gain: float = 0.1,
                   threshold: float = 0.0, clean_dtype=torch.float32):
        """Hogbom CLEAN of every plane of the dirty image (channel psf, first
        polarisation); returns (components, residual images, iterations per
        plane), components and residuals in float64."""
        nch, npol = dirty.shape[:2]
        if self._work is None:
            self._work = H._work(self.dev)
        comps = torch.zeros(dirty.shape, dtype=torch.float64, device=self.dev)
        resid = torch.empty(dirty.shape, dtype=torch.float64, device=self.dev)
        iters = np.zeros((nch, npol), dtype=np.int64)
        for c in range(nch):
            p = psf[c, 0].to(clean_dtype).contiguous()
            for q in range(npol):
                r = dirty[c, q].to(clean_dtype).contiguous()
                m = torch.zeros_like(r)
                iters[c, q] = H.clean_device(r, p, m, None, None, int(niter), 0, gain, threshold,
                                             work=self._work)
                comps[c, q] = m.to(torch.float64)
                resid[c, q] = r.to(torch.float64)
        return comps, resid, iters

    def write_back(self, ms4) -> None:
        """Copy the model and residual visibilities into the dataset's
        VISIBILITY_MODEL and RESIDUAL (complex128, the VISIBILITY dims)."""
        from .msv4 import make_data_array
        v = ms4["VISIBILITY"]
        for name, t in (("VISIBILITY_MODEL", self.model_vis), ("RESIDUAL", self.residual_vis)):
            ms4[name] = make_data_array(t.cpu().numpy().astype(np.complex128).reshape(self.shape),
                                        coords=v.coords, dims=v.dims)

    def close(self) -> None:
        self.plan.close()


def run_major_cycles(ms4, image_size, cell_size, nmajor: int = 3, cycle_niter: int = 100,
                     gain: float = 0.1, threshold: float = 0.0, support: int = 7,
                     oversampling: int = 100, chan_mode: str = "cube", precision: str = "fp64",
                     initial_model: Optional[np.ndarray] = None, device=None):
    """Cotton-Schwab major cycles on device-resident data: each cycle grids
    the residual visibilities, CLEANs cycle_niter iterations per plane down
    to `threshold`, adds the components to the model, predicts the model
    visibilities and forms the residual ones.  The loop ends after nmajor
    cycles or when no plane's peak residual exceeds the threshold.  Returns
    (model [n_img_chan, n_pol, ny, nx] float64, last residual image, PSF,
    per-cycle log).  This is the major/minor structure of
    imager.run_imaging_loop without its IterationController heuristics
    (cycle thresholds from PSF sidelobes) or Stokes conversion."""
    im = ResidentImager(ms4, image_size, cell_size, support, oversampling, chan_mode, precision,
                        device=device)
    try:
        psf = im.make_psf()
        model = torch.zeros((im.n_img_chan, im.shape[3], im.ny, im.nx), dtype=torch.float64,
                            device=im.dev)
        if initial_model is not None:
            model += torch.as_tensor(initial_model, dtype=torch.float64).to(im.dev)
            im.predict_model_visibilities(model)
            im.compute_residual_visibilities()
        log = []
        residual = None
        for cycle in range(nmajor):
            dirty = im.grid_visibilities("RESIDUAL")
            peak = float(dirty.abs().max())
            if peak <= threshold:
                residual = dirty
                log.append({"cycle": cycle, "peak_before": peak, "iterations": 0})
                break
            comps, residual, iters = im.deconvolve(dirty, psf, cycle_niter, gain, threshold)
            model += comps
            im.predict_model_visibilities(model)
            im.compute_residual_visibilities()
            log.append({"cycle": cycle, "peak_before": peak,
                        "peak_after": float(residual.abs().max()),
                        "iterations": int(iters.sum())})
        return model, residual, psf, log
    finally:
        im.close()
