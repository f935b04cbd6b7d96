"""Seeded synthetic interferometer data shaped like an MSv4 partition.

Array: antennas uniform in a disc; uvw by earth rotation of the baselines
(Thompson, Moran & Swenson eq. 4.1) at a fixed declination over a range of
hour angles; visibilities are the exact DFT of point sources plus complex
Gaussian noise.  Small cases are built with NumPy on the host; the large
benchmark cases are built directly in GPU memory with torch.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

C_LIGHT = 299792458.0
ARCSEC = math.pi / 180.0 / 3600.0
VLA_LATITUDE_DEG = 34.0784


@dataclass
class ArrayConfig:
    n_ant: int = 27
    radius_m: float = 20e3
    n_time: int = 100
    ha_range_h: tuple = (-2.0, 2.0)
    dec_deg: float = 30.0
    lat_deg: float = VLA_LATITUDE_DEG
    freq_lo: float = 1.0e9
    freq_hi: float = 1.4e9
    n_chan: int = 4
    n_pol: int = 2
    seed: int = 1


@dataclass
class SkyConfig:
    n_src: int = 20
    flux: tuple = (0.1, 1.0)
    max_offset_deg: float = 0.5
    noise: float = 0.01


# the configurations of the benchmark/test suite
CONFIG1 = dict(array=ArrayConfig(), sky=SkyConfig(), n_uv=(256, 256), cell_arcsec=5.0,
               flag_frac=0.0, weight_range=None)
CONFIG2 = dict(array=ArrayConfig(n_ant=64, n_time=1000, n_chan=64, n_pol=4, seed=2),
               sky=SkyConfig(), n_uv=(2048, 2048), cell_arcsec=1.0, flag_frac=0.1,
               weight_range=(0.5, 1.5))


def antenna_enu(n_ant: int, radius_m: float, rng: np.random.Generator) -> np.ndarray:
    r = radius_m * np.sqrt(rng.random(n_ant))
    th = 2.0 * np.pi * rng.random(n_ant)
    return np.stack([r * np.cos(th), r * np.sin(th), np.zeros(n_ant)], axis=1)


def baseline_pairs(n_ant: int) -> np.ndarray:
    i, j = np.triu_indices(n_ant, k=1)
    return np.stack([i, j], axis=1)


def uvw_tracks(enu: np.ndarray, pairs: np.ndarray, hour_angles: np.ndarray,
               dec_rad: float, lat_rad: float) -> np.ndarray:
    """uvw [n_time, n_baseline, 3] in metres."""
    e, n, u_ = enu[:, 0], enu[:, 1], enu[:, 2]
    x = -math.sin(lat_rad) * n + math.cos(lat_rad) * u_
    y = e
    z = math.cos(lat_rad) * n + math.sin(lat_rad) * u_
    lx = x[pairs[:, 1]] - x[pairs[:, 0]]
    ly = y[pairs[:, 1]] - y[pairs[:, 0]]
    lz = z[pairs[:, 1]] - z[pairs[:, 0]]
    h = hour_angles[:, None]
    sd, cd = math.sin(dec_rad), math.cos(dec_rad)
    uu = np.sin(h) * lx + np.cos(h) * ly
    vv = -sd * np.cos(h) * lx + sd * np.sin(h) * ly + cd * lz
    ww = cd * np.cos(h) * lx - cd * np.sin(h) * ly + sd * lz
    return np.stack([uu, vv, ww], axis=2)


@dataclass
class Observation:
    uvw: np.ndarray            # [nt, nb, 3] float64
    freq: np.ndarray           # [nchan] float64
    src_lm: np.ndarray         # [nsrc, 2] radians
    src_flux: np.ndarray       # [nsrc]
    n_pol: int
    seed: int
    meta: dict = field(default_factory=dict)


def make_observation(array: ArrayConfig, sky: SkyConfig, n_time: Optional[int] = None
                     ) -> Observation:
    rng = np.random.default_rng(array.seed)
    enu = antenna_enu(array.n_ant, array.radius_m, rng)
    pairs = baseline_pairs(array.n_ant)
    nt = array.n_time if n_time is None else n_time
    ha = np.linspace(array.ha_range_h[0], array.ha_range_h[1], nt) * (np.pi / 12.0)
    uvw = uvw_tracks(enu, pairs, ha, math.radians(array.dec_deg), math.radians(array.lat_deg))
    freq = np.linspace(array.freq_lo, array.freq_hi, array.n_chan)
    off = math.radians(sky.max_offset_deg)
    lm = rng.uniform(-off, off, size=(sky.n_src, 2))
    flux = rng.uniform(sky.flux[0], sky.flux[1], size=sky.n_src)
    return Observation(uvw=uvw, freq=freq, src_lm=lm, src_flux=flux, n_pol=array.n_pol,
                       seed=array.seed,
                       meta=dict(n_ant=array.n_ant, radius_m=array.radius_m,
                                 dec_deg=array.dec_deg, lat_deg=array.lat_deg,
                                 ha_range_h=array.ha_range_h))


def host_visibilities(obs: Observation, sky: SkyConfig, flag_frac: float = 0.0,
                      weight_range=None, rng_seed: Optional[int] = None):
    """NumPy complex128 vis, float64 weight, bool flag [nt, nb, nchan, npol]."""
    rng = np.random.default_rng(obs.seed + 1000 if rng_seed is None else rng_seed)
    nt, nb, _ = obs.uvw.shape
    nf, npol = obs.freq.size, obs.n_pol
    vis = np.zeros((nt, nb, nf), dtype=np.complex128)
    for (l, m), s in zip(obs.src_lm, obs.src_flux):
        ph = obs.uvw[:, :, 0:1] * l + obs.uvw[:, :, 1:2] * m  # metres * rad
        vis += s * np.exp(-2j * np.pi * ph * (obs.freq[None, None, :] / C_LIGHT))
    vis = np.repeat(vis[..., None], npol, axis=3)
    vis = vis + sky.noise * (rng.standard_normal(vis.shape) + 1j * rng.standard_normal(vis.shape))
    if weight_range is None:
        w = np.ones(vis.shape, dtype=np.float64)
    else:
        w = rng.uniform(weight_range[0], weight_range[1], size=vis.shape)
    flag = rng.random(vis.shape) < flag_frac if flag_frac > 0 else np.zeros(vis.shape, bool)
    return vis, w, flag


def device_visibilities(obs: Observation, sky: SkyConfig, flag_frac: float = 0.0,
                        weight_range=None, device="cuda", chunk_chan: int = 4):
    """Device complex64 vis, float32 weight, uint8 flag [rows, nchan, npol],
    built with torch on the GPU (seeded), plus uvw [rows, 3] float64."""
    import torch

    nt, nb, _ = obs.uvw.shape
    rows = nt * nb
    nf, npol = obs.freq.size, obs.n_pol
    dev = torch.device(device)
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(obs.seed) * 7919 + 17)
    uvw = torch.from_numpy(np.ascontiguousarray(obs.uvw.reshape(rows, 3))).to(dev)
    vis = torch.empty((rows, nf, npol), dtype=torch.complex64, device=dev)
    u = uvw[:, 0:1]
    v = uvw[:, 1:2]
    for c0 in range(0, nf, chunk_chan):
        c1 = min(nf, c0 + chunk_chan)
        f = torch.from_numpy(obs.freq[c0:c1]).to(dev)[None, :] / C_LIGHT
        acc = torch.zeros((rows, c1 - c0), dtype=torch.complex128, device=dev)
        for (l, m), s in zip(obs.src_lm, obs.src_flux):
            ph = (-2.0 * math.pi) * (u * float(l) + v * float(m)) * f
            acc += float(s) * torch.polar(torch.ones_like(ph), ph)
        blk = acc.to(torch.complex64)[:, :, None].expand(rows, c1 - c0, npol).clone()
        noise = torch.randn((rows, c1 - c0, npol, 2), generator=gen, device=dev,
                            dtype=torch.float32) * sky.noise
        blk += torch.view_as_complex(noise)
        vis[:, c0:c1, :] = blk
        del acc, blk, noise
    if weight_range is None:
        w = torch.ones((rows, nf, npol), dtype=torch.float32, device=dev)
    else:
        w = torch.empty((rows, nf, npol), dtype=torch.float32, device=dev)
        w.uniform_(weight_range[0], weight_range[1], generator=gen)
    if flag_frac > 0:
        flag = (torch.rand((rows, nf, npol), generator=gen, device=dev) < flag_frac)
        flag = flag.to(torch.uint8)
    else:
        flag = torch.zeros((rows, nf, npol), dtype=torch.uint8, device=dev)
    return uvw, vis, w, flag


def uneven_plane_sets(n_planes: int = 256, n_uv: int = 512, counts=(1e4, 2e6),
                      flag_range=(0.0, 0.6), supports=(5, 7, 9), n_pol: int = 4,
                      seed: int = 4, device="cuda", array: Optional[ArrayConfig] = None,
                      sky: Optional[SkyConfig] = None, dtype=None, chunk: int = 1 << 24):
    """Batched uneven workload: `n_planes` (channel, polarisation) planes of
    n_uv x n_uv.  Plane p belongs to channel p // n_pol and uses support
    supports[p % len(supports)]; its unflagged count is log-uniform in
    `counts`, and a fraction U(flag_range) of its samples is flagged on top.
    Each sample is one baseline (64-antenna array) at a random hour angle in
    the array's range; visibilities are the point-source DFT plus noise,
    weights U(0.5, 1.5).  The cell puts the longest baseline at the highest
    frequency at 0.95 of the grid's half width.

    Returns ([PlaneSet per support], (cell_u, cell_v), info)."""
    import torch

    from .batch import PlaneSet

    arr = array or ArrayConfig(n_ant=64, seed=seed)
    sky = sky or SkyConfig()
    rng = np.random.default_rng(seed)
    dev = torch.device(device)
    cdtype = dtype or torch.complex64
    rdtype = torch.float32 if cdtype == torch.complex64 else torch.float64
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed) * 104729 + 3)

    enu = antenna_enu(arr.n_ant, arr.radius_m, rng)
    pairs = baseline_pairs(arr.n_ant)
    lat = math.radians(arr.lat_deg)
    e, n, up = enu[:, 0], enu[:, 1], enu[:, 2]
    x = -math.sin(lat) * n + math.cos(lat) * up
    y = e
    z = math.cos(lat) * n + math.sin(lat) * up
    bl = np.stack([x[pairs[:, 1]] - x[pairs[:, 0]], y[pairs[:, 1]] - y[pairs[:, 0]],
                   z[pairs[:, 1]] - z[pairs[:, 0]]], axis=1)
    n_chan = (n_planes + n_pol - 1) // n_pol
    freqs = np.linspace(arr.freq_lo, arr.freq_hi, n_chan)
    plane_freq = freqs[np.arange(n_planes) // n_pol]
    off = math.radians(sky.max_offset_deg)
    lm = rng.uniform(-off, off, size=(sky.n_src, 2))
    flux = rng.uniform(sky.flux[0], sky.flux[1], size=sky.n_src)
    fflag = rng.uniform(flag_range[0], flag_range[1], size=n_planes)
    n_after = 10.0 ** rng.uniform(math.log10(counts[0]), math.log10(counts[1]), size=n_planes)
    n_tot = np.maximum(1, np.round(n_after / (1.0 - fflag))).astype(np.int64)
    bmax = float(np.max(np.linalg.norm(bl, axis=1)))
    cell = 0.95 / (2.0 * bmax * arr.freq_hi / C_LIGHT)
    dec = math.radians(arr.dec_deg)
    sd, cd = math.sin(dec), math.cos(dec)
    ha0, ha1 = (h * math.pi / 12.0 for h in arr.ha_range_h)
    blt = torch.from_numpy(bl).to(dev)

    sets = []
    unflagged = 0
    for s in sorted(set(supports)):
        planes = [p for p in range(n_planes) if supports[p % len(supports)] == s]
        cnt = torch.from_numpy(n_tot[planes]).to(dev)
        total = int(cnt.sum().item())
        plane = torch.repeat_interleave(torch.arange(len(planes), device=dev, dtype=torch.int32),
                                        cnt)
        pf = torch.from_numpy(plane_freq[planes]).to(dev)
        pflag = torch.from_numpy(fflag[planes]).to(dev)
        uvw = torch.empty((total, 3), dtype=torch.float64, device=dev)
        vis = torch.empty((total, 1), dtype=cdtype, device=dev)
        flag = torch.empty((total, 1), dtype=torch.uint8, device=dev)
        for a in range(0, total, chunk):
            b = min(total, a + chunk)
            k = torch.randint(0, bl.shape[0], (b - a,), generator=gen, device=dev)
            h = ha0 + (ha1 - ha0) * torch.rand((b - a,), generator=gen, device=dev,
                                               dtype=torch.float64)
            lx, ly, lz = blt[k, 0], blt[k, 1], blt[k, 2]
            sh, ch = torch.sin(h), torch.cos(h)
            uu = sh * lx + ch * ly
            vv = -sd * ch * lx + sd * sh * ly + cd * lz
            ww = cd * ch * lx - cd * sh * ly + sd * lz
            uvw[a:b] = torch.stack([uu, vv, ww], dim=1)
            f = pf[plane[a:b].long()] / C_LIGHT
            acc = torch.zeros((b - a,), dtype=torch.complex128, device=dev)
            for (l, m), fl in zip(lm, flux):
                ph = (-2.0 * math.pi) * (uu * float(l) + vv * float(m)) * f
                acc += float(fl) * torch.polar(torch.ones_like(ph), ph)
            noise = torch.randn((b - a, 2), generator=gen, device=dev,
                                dtype=torch.float64) * sky.noise
            acc += torch.view_as_complex(noise)
            vis[a:b, 0] = acc.to(cdtype)
            fr = torch.rand((b - a,), generator=gen, device=dev, dtype=torch.float64)
            flag[a:b, 0] = (fr < pflag[plane[a:b].long()]).to(torch.uint8)
        weight = torch.empty((total, 1), dtype=rdtype, device=dev)
        weight.uniform_(0.5, 1.5, generator=gen)
        unflagged += total - int(flag.sum().item())
        sets.append(PlaneSet(support=s, uvw=uvw, plane=plane, freq=pf, vis=vis, weight=weight,
                             flag=flag))
    info = dict(unflagged=unflagged, planes=n_planes, n_uv=n_uv, seed=seed,
                plane_counts=n_tot.tolist())
    return sets, (-cell, cell), info


def blob_sky(n: int, nblob: int, rng: np.random.Generator, sigma=(2.0, 10.0),
             amp=(0.1, 1.0)) -> np.ndarray:
    """n x n image of `nblob` Gaussian blobs, widths U(sigma) pixels,
    amplitudes U(amp), centres uniform over the image."""
    yy, xx = np.mgrid[0:n, 0:n]
    img = np.zeros((n, n))
    for _ in range(nblob):
        cx, cy = rng.uniform(0, n, size=2)
        s = rng.uniform(*sigma)
        a = rng.uniform(*amp)
        r0 = int(max(0, cy - 6 * s)), int(min(n, cy + 6 * s + 1))
        c0 = int(max(0, cx - 6 * s)), int(min(n, cx + 6 * s + 1))
        sub = (slice(r0[0], r0[1]), slice(c0[0], c0[1]))
        img[sub] += a * np.exp(-((xx[sub] - cx) ** 2 + (yy[sub] - cy) ** 2) / (2 * s * s))
    return img


def measurement_samples(points: int, n: int, seed: int = 3):
    """About `points` uvw samples for the measurement operator: the
    27-antenna, 4-channel array (351 baselines) with its hour-angle range
    sampled finely enough to reach `points` (rows x channels), truncated to
    exactly `points` // 4 rows.  The cell puts the longest baseline at the
    highest frequency at 0.45 of the grid's half width.  Returns
    (uvw [rows, 3], freq [4], (cell_u, cell_v))."""
    arr = ArrayConfig(seed=seed)
    nb = arr.n_ant * (arr.n_ant - 1) // 2
    rows = points // arr.n_chan
    nt = (rows + nb - 1) // nb
    obs = make_observation(ArrayConfig(**{**arr.__dict__, "n_time": nt}), SkyConfig())
    uvw = np.ascontiguousarray(obs.uvw.reshape(-1, 3)[:rows])
    maxuv = np.max(np.hypot(uvw[:, 0], uvw[:, 1]))
    c = 0.45 * C_LIGHT / obs.freq[-1] / maxuv
    return uvw, obs.freq, (-c, c)
