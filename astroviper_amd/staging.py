"""Host -> GPU staging of MSv4 partitions for the gridder.

The reference reads a processing set partition by partition
(make_image_single_field.py:73-89, xradio's ProcessingSetIterator) and, for
each, zeroes WEIGHT where FLAG is set or UVW is NaN and grids the partition
on the host (standard_grid.py:660-670).  Here a partition's rows (time x
baseline) are streamed to the GPU in blocks:

* block k is read from its source (NumPy arrays, memory maps or zarr chunk
  files) into pinned host slot k % 2 by the CPU;
* a copy stream moves the slot to the device (no host synchronisation);
* the compute stream waits for that copy, bins the block's uvw with one plan
  reused for every block (a partial last block pads its uvw with NaN rows,
  which the plan drops) and grids it into the accumulating grid;
* the CPU refills a slot only after the compute of the block that last used
  it has finished, so the read of block k + 1 overlaps the copy and the
  gridding of block k.

Flags and NaN uvw are applied on the GPU: the gridder zeroes the weight of a
flagged polarisation and the plan drops a sample whose uvw is NaN, which is
what the reference's host-side WEIGHT masking achieves.  Device memory is
bounded by the block size, not by the partition.

On-disk partitions are zarr v2 stores (the MSv4 storage format): a
directory holding the arrays VISIBILITY (time, baseline, frequency,
polarization), UVW (time, baseline, 3), WEIGHT, FLAG and frequency, each a
directory with a .zarray header and chunk files.  zarr itself is not
installed here, so ZarrV2Array reads the format directly; chunks stored raw
or with the zlib or gzip codec are read, the Blosc and Zstd codecs (which
need compiled libraries this environment lacks) are refused with an error.
"""
from __future__ import annotations

import gzip
import json
import os
import zlib
from typing import Dict, Iterator, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import device as _dev

ARRAYS = ("VISIBILITY", "UVW", "WEIGHT", "FLAG")


# ------------------------------------------------------------ zarr v2 I/O
class ZarrV2Array:
    """Read access to one zarr v2 array directory (C order)."""

    def __init__(self, path: str):
        self.path = path
        with open(os.path.join(path, ".zarray")) as f:
            meta = json.load(f)
        if meta.get("zarr_format") != 2:
            raise ValueError(f"{path}: not a zarr v2 array")
        if meta.get("order", "C") != "C":
            raise NotImplementedError(f"{path}: only C-order chunks are read")
        if meta.get("filters"):
            raise NotImplementedError(f"{path}: zarr filters are not supported")
        comp = meta.get("compressor")
        self.codec = None if comp is None else comp.get("id")
        if self.codec not in (None, "zlib", "gzip"):
            raise NotImplementedError(
                f"{path}: chunk codec {self.codec!r} needs a compression library that is not "
                "installed here (raw, zlib and gzip chunks are read)")
        self.shape = tuple(int(s) for s in meta["shape"])
        self.chunks = tuple(int(c) for c in meta["chunks"])
        self.dtype = np.dtype(meta["dtype"])
        self.fill_value = meta.get("fill_value")
        self.sep = meta.get("dimension_separator", ".")

    @property
    def ndim(self) -> int:
        return len(self.shape)

    def _chunk(self, idx: Tuple[int, ...]) -> Optional[np.ndarray]:
        name = self.sep.join(str(i) for i in idx) if idx else "0"
        fn = os.path.join(self.path, name)
        if not os.path.exists(fn):
            return None
        with open(fn, "rb") as f:
            raw = f.read()
        if self.codec == "zlib":
            raw = zlib.decompress(raw)
        elif self.codec == "gzip":
            raw = gzip.decompress(raw)
        return np.frombuffer(raw, dtype=self.dtype).reshape(self.chunks)

    def read(self, a0: int = 0, a1: Optional[int] = None) -> np.ndarray:
        """The sub-array [a0:a1] along axis 0, all of the other axes."""
        a1 = self.shape[0] if a1 is None else min(a1, self.shape[0])
        out_shape = (max(0, a1 - a0),) + self.shape[1:]
        fill = 0 if self.fill_value is None else self.fill_value
        out = np.full(out_shape, fill, dtype=self.dtype)
        if out.size == 0:
            return out
        c0 = self.chunks[0]
        grids = [range(-(-s // c)) for s, c in zip(self.shape[1:], self.chunks[1:])]
        for i0 in range(a0 // c0, (a1 - 1) // c0 + 1):
            for rest in np.ndindex(*[len(g) for g in grids]) if grids else [()]:
                idx = (i0,) + tuple(rest)
                blk = self._chunk(idx)
                if blk is None:
                    continue
                src, dst = [], []
                lo0, hi0 = i0 * c0, min((i0 + 1) * c0, self.shape[0])
                s0, e0 = max(lo0, a0), min(hi0, a1)
                src.append(slice(s0 - lo0, e0 - lo0))
                dst.append(slice(s0 - a0, e0 - a0))
                for ax, j in enumerate(rest, start=1):
                    c = self.chunks[ax]
                    lo, hi = j * c, min((j + 1) * c, self.shape[ax])
                    src.append(slice(0, hi - lo))
                    dst.append(slice(lo, hi))
                out[tuple(dst)] = blk[tuple(src)]
        return out


def write_zarr_v2(path: str, data: np.ndarray, chunks: Sequence[int],
                  compressor: Optional[str] = None) -> None:
    """Write `data` as a zarr v2 array (raw, zlib or gzip chunks, C order)."""
    data = np.ascontiguousarray(data)
    os.makedirs(path, exist_ok=True)
    chunks = tuple(int(min(c, s)) if s else 1 for c, s in zip(chunks, data.shape))
    meta = {"zarr_format": 2, "shape": list(data.shape), "chunks": list(chunks),
            "dtype": data.dtype.str, "order": "C", "filters": None, "fill_value": 0,
            "compressor": None if compressor is None else {"id": compressor, "level": 1}}
    if data.dtype == np.bool_:
        meta["fill_value"] = False
    with open(os.path.join(path, ".zarray"), "w") as f:
        json.dump(meta, f)
    grid = [-(-s // c) for s, c in zip(data.shape, chunks)]
    for idx in np.ndindex(*grid):
        sl = tuple(slice(i * c, min((i + 1) * c, s)) for i, c, s in zip(idx, chunks, data.shape))
        blk = np.zeros(chunks, dtype=data.dtype)  # edge chunks are stored full size
        part = data[sl]
        blk[tuple(slice(0, n) for n in part.shape)] = part
        raw = blk.tobytes()
        if compressor == "zlib":
            raw = zlib.compress(raw, 1)
        elif compressor == "gzip":
            raw = gzip.compress(raw, 1)
        with open(os.path.join(path, ".".join(str(i) for i in idx)), "wb") as f:
            f.write(raw)


# ------------------------------------------------------------ partitions
class Partition:
    """One MSv4 partition as row blocks: rows are (time, baseline) pairs,
    read whole time steps at a time."""

    def __init__(self, read, n_time: int, n_baseline: int, freq: np.ndarray, n_pol: int,
                 vis_dtype, weight_dtype):
        self._read = read
        self.n_time, self.n_baseline = int(n_time), int(n_baseline)
        self.freq = np.asarray(freq, dtype=np.float64)
        self.n_chan, self.n_pol = self.freq.size, int(n_pol)
        self.vis_dtype, self.weight_dtype = np.dtype(vis_dtype), np.dtype(weight_dtype)

    @property
    def n_rows(self) -> int:
        return self.n_time * self.n_baseline

    def read_times(self, t0: int, t1: int) -> Dict[str, np.ndarray]:
        """VISIBILITY, UVW, WEIGHT, FLAG of time steps [t0, t1)."""
        return self._read(t0, t1)


def array_partition(vis, uvw, weight, flag, freq) -> Partition:
    """A partition over arrays [n_time, n_baseline, ...] (vis may be None
    for weights only; flag may be None)."""
    weight = np.asarray(weight)
    nt, nb, nchan, npol = weight.shape

    def read(t0, t1):
        return {"VISIBILITY": None if vis is None else vis[t0:t1], "UVW": uvw[t0:t1],
                "WEIGHT": weight[t0:t1],
                "FLAG": (np.zeros((t1 - t0,) + weight.shape[1:], np.uint8) if flag is None
                         else flag[t0:t1])}

    vdt = np.complex128 if vis is None else np.asarray(vis[:0]).dtype
    return Partition(read, nt, nb, freq, npol, vdt, weight.dtype)


def memory_partition(ds, column: str = "VISIBILITY") -> Partition:
    """A partition over in-memory arrays (an MSv4 dataset: xarray or
    msv4.VisibilityDataset)."""
    vis, uvw = ds[column].data, ds.UVW.data
    w, fl = ds.WEIGHT.data, ds.FLAG.data
    freq = ds.coords["frequency"].values
    nt, nb, nchan, npol = w.shape

    def read(t0, t1):
        return {"VISIBILITY": vis[t0:t1], "UVW": uvw[t0:t1], "WEIGHT": w[t0:t1],
                "FLAG": fl[t0:t1]}

    return Partition(read, nt, nb, freq, npol, vis.dtype, w.dtype)


def _find_arrays(path: str) -> Optional[str]:
    """The directory under `path` (itself or one level down) holding the
    VISIBILITY, UVW, WEIGHT and FLAG arrays."""
    def has(d):
        return all(os.path.exists(os.path.join(d, a, ".zarray")) for a in ARRAYS)
    if has(path):
        return path
    for sub in sorted(os.listdir(path)):
        d = os.path.join(path, sub)
        if os.path.isdir(d) and has(d):
            return d
    return None


def zarr_partition(path: str, column: str = "VISIBILITY") -> Partition:
    """A partition stored as zarr v2 arrays (see the module notes)."""
    root = _find_arrays(path)
    if root is None:
        raise FileNotFoundError(f"{path}: no VISIBILITY/UVW/WEIGHT/FLAG zarr arrays")
    arr = {a: ZarrV2Array(os.path.join(root, column if a == "VISIBILITY" else a))
           for a in ARRAYS}
    fpath = os.path.join(root, "frequency")
    if not os.path.exists(os.path.join(fpath, ".zarray")):
        fpath = os.path.join(path, "frequency")
    freq = ZarrV2Array(fpath).read()
    nt, nb, nchan, npol = arr["WEIGHT"].shape

    def read(t0, t1):
        return {a: arr[a].read(t0, t1) for a in ARRAYS}

    return Partition(read, nt, nb, freq, npol, arr["VISIBILITY"].dtype, arr["WEIGHT"].dtype)


def open_processing_set(path: str, column: str = "VISIBILITY") -> List[Partition]:
    """Every partition of a processing-set directory (or the one partition
    `path` itself holds), in name order."""
    if _find_arrays(path) == path:
        return [zarr_partition(path, column)]
    parts = []
    for sub in sorted(os.listdir(path)):
        d = os.path.join(path, sub)
        if os.path.isdir(d) and _find_arrays(d) is not None:
            parts.append(zarr_partition(d, column))
    if not parts:
        raise FileNotFoundError(f"{path}: no partitions")
    return parts


def write_partition(path: str, vis, uvw, weight, flag, freq, time_chunk: int = 16,
                    compressor: Optional[str] = None) -> None:
    """Write one partition as zarr v2 arrays chunked along time."""
    nt, nb, nchan, npol = np.shape(weight)
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, ".zgroup"), "w") as f:
        json.dump({"zarr_format": 2}, f)
    write_zarr_v2(os.path.join(path, "VISIBILITY"), vis, (time_chunk, nb, nchan, npol),
                  compressor)
    write_zarr_v2(os.path.join(path, "UVW"), uvw, (time_chunk, nb, 3), compressor)
    write_zarr_v2(os.path.join(path, "WEIGHT"), weight, (time_chunk, nb, nchan, npol),
                  compressor)
    write_zarr_v2(os.path.join(path, "FLAG"), flag, (time_chunk, nb, nchan, npol), compressor)
    write_zarr_v2(os.path.join(path, "frequency"), np.asarray(freq, np.float64), (nchan,))


# ------------------------------------------------------------ the stager
def _host_tensor(a) -> torch.Tensor:
    """A CPU tensor over a host array (no copy for NumPy arrays; booleans
    as uint8)."""
    a = np.asarray(a)
    if a.dtype == np.bool_:
        a = a.view(np.uint8)
    if not a.flags.c_contiguous:
        a = np.ascontiguousarray(a)
    return torch.from_numpy(a)


class StagedGridder:
    """Grid partitions streamed through two pinned host slots (see the
    module notes) into one device grid [n_img_chan, n_pol, n_u, n_v].

    ``fp32`` grids in FP32 (complex64 visibilities on the device), else in
    FP64; coordinates and sums of weights are FP64 either way.
    ``block_rows`` bounds the rows of a block (whole time steps of the
    partition, at least one).
    """

    def __init__(self, n_uv, cell, n_chan: int, n_pol: int, chan_mode: str = "continuum",
                 support: int = 7, oversampling: int = 100, fp32: bool = True,
                 block_rows: int = 1 << 18, tile=(16, 16), device=None, psf: bool = False,
                 cgk: Optional[torch.Tensor] = None, geometry: str = "reference",
                 chan_map=None, n_img_chan: Optional[int] = None,
                 grid: Optional[torch.Tensor] = None, sumwt: Optional[torch.Tensor] = None):
        self.device = _dev.require_gpu(device)
        self.n_uv = (int(n_uv[0]), int(n_uv[1]))
        self.cell = (float(cell[0]), float(cell[1]))
        self.n_chan, self.n_pol = int(n_chan), int(n_pol)
        if chan_map is None:
            cube = chan_mode == "cube"
            self.n_img_chan = self.n_chan if cube else 1
            chan_map = np.arange(self.n_chan) if cube else np.zeros(self.n_chan, dtype=int)
        else:
            self.n_img_chan = int(n_img_chan)
        self.chan_map = torch.from_numpy(np.ascontiguousarray(chan_map, dtype=np.int32)).to(
            self.device)
        self.support, self.oversampling = int(support), int(oversampling)
        self.fp32, self.psf = bool(fp32), bool(psf)
        self.block_rows, self.tile, self.geometry = int(block_rows), tuple(tile), geometry
        self.cdt = torch.complex64 if self.fp32 else torch.complex128
        self.rdt = torch.float32 if self.fp32 else torch.float64
        shape = (self.n_img_chan, self.n_pol) + self.n_uv
        self.grid = (torch.zeros(shape, dtype=self.cdt, device=self.device) if grid is None
                     else grid)
        self.sumwt = (torch.zeros(shape[:2], dtype=torch.float64, device=self.device)
                      if sumwt is None else sumwt)
        if tuple(self.grid.shape) != shape or self.grid.dtype != self.cdt:
            raise ValueError(f"grid must be {shape} {self.cdt}")
        self.cgk = cgk if cgk is not None else _dev.ps_kernel_1d(self.oversampling,
                                                                 self.support, self.device)
        self.copy_stream = torch.cuda.Stream(device=self.device)
        self._slots = None
        self._plan = None
        self.rows_gridded = 0
        self.blocks = 0

    # one set of pinned host and device buffers per slot
    def _alloc(self, rows: int, nb: int):
        nc, npol = self.n_chan, self.n_pol
        dev = self.device
        vdt_h = np.complex64 if self.fp32 else np.complex128
        wdt_h = np.float32 if self.fp32 else np.float64
        slots = []
        for _ in range(2):
            host = {"UVW": torch.empty((rows, 3), dtype=torch.float64, pin_memory=True),
                    "WEIGHT": torch.empty((rows, nc, npol), dtype=self.rdt, pin_memory=True),
                    "FLAG": torch.empty((rows, nc, npol), dtype=torch.uint8, pin_memory=True)}
            devb = {"UVW": torch.empty((rows, 3), dtype=torch.float64, device=dev),
                    "WEIGHT": torch.empty((rows, nc, npol), dtype=self.rdt, device=dev),
                    "FLAG": torch.empty((rows, nc, npol), dtype=torch.uint8, device=dev)}
            if not self.psf:
                host["VISIBILITY"] = torch.empty((rows, nc, npol), dtype=self.cdt,
                                                 pin_memory=True)
                devb["VISIBILITY"] = torch.empty((rows, nc, npol), dtype=self.cdt, device=dev)
            slots.append({"host": host, "dev": devb, "copied": torch.cuda.Event(),
                          "done": torch.cuda.Event(), "used": False})
        self._slots = slots
        self._slot_rows = rows
        self._vdt_h, self._wdt_h = vdt_h, wdt_h
        self._plan = _dev.GridPlan(rows, self.n_chan, self.n_img_chan, self.n_uv, self.cell,
                                   self.support, self.oversampling, tile=self.tile,
                                   geometry=self.geometry, device=dev, n_baseline=nb)
        self._slot_nb = nb
        self._plan.cgk = self.cgk
        self._freq = None

    def add(self, part: Partition) -> None:
        """Grid one partition (accumulating)."""
        if part.n_chan != self.n_chan or part.n_pol < self.n_pol:
            raise ValueError("partition channels/polarisations do not match the grid")
        nb = part.n_baseline
        t_per = max(1, self.block_rows // max(1, nb))
        rows = t_per * nb
        if self._slots is None or self._slot_rows != rows or self._slot_nb != nb:
            self._alloc(rows, nb)
        freq_t = torch.from_numpy(np.ascontiguousarray(part.freq)).to(self.device)
        cur = torch.cuda.current_stream(self.device)
        k = 0
        for t0 in range(0, part.n_time, t_per):
            t1 = min(part.n_time, t0 + t_per)
            n = (t1 - t0) * nb
            slot = self._slots[k % 2]
            if slot["used"]:
                slot["done"].synchronize()  # the compute that last read this slot is over
            blk = part.read_times(t0, t1)
            h = slot["host"]
            # into the pinned slot with torch's (multithreaded) CPU copies,
            # converting to the slot's types on the way
            h["UVW"][:n].copy_(_host_tensor(blk["UVW"]).reshape(n, 3))
            for name in (("WEIGHT", "FLAG") if self.psf else ("WEIGHT", "FLAG", "VISIBILITY")):
                src = _host_tensor(blk[name]).reshape(n, self.n_chan, -1)[:, :, :self.n_pol]
                h[name][:n].copy_(src)
            with torch.cuda.stream(self.copy_stream):
                if slot["used"]:
                    self.copy_stream.wait_event(slot["done"])
                for name, ht in slot["host"].items():
                    slot["dev"][name][:n].copy_(ht[:n], non_blocking=True)
                slot["copied"].record(self.copy_stream)
            cur.wait_event(slot["copied"])
            d = slot["dev"]
            if n < rows:  # a partial block: its stale rows become NaN uvw, dropped
                d["UVW"][n:].fill_(float("nan"))
            self._plan.set_uvw(d["UVW"], freq_t, self.chan_map)
            self._plan.grid(None if self.psf else d["VISIBILITY"], d["WEIGHT"], d["FLAG"],
                            grid=self.grid, sumwt=self.sumwt, n_pol=self.n_pol, psf=self.psf,
                            accumulate=True, dtype=self.rdt)
            slot["done"].record(cur)
            slot["used"] = True
            self.rows_gridded += n
            self.blocks += 1
            k += 1

    def close(self) -> None:
        if self._slots is not None:
            for s in self._slots:
                if s["used"]:
                    s["done"].synchronize()
        if self._plan is not None:
            self._plan.close()
            self._plan = None
        self._slots = None


def grid_processing_set(partitions, n_uv, cell, chan_mode: str = "continuum",
                        dopsf: bool = False, fp32: bool = True, support: int = 7,
                        oversampling: int = 100, block_rows: int = 1 << 18,
                        column: str = "VISIBILITY"):
    """Grid every partition of a processing set (a directory, a list of
    directories, datasets or Partitions) into one device grid; returns
    (grid, sumwt) on the GPU."""
    if isinstance(partitions, str):
        parts = open_processing_set(partitions, column)
    else:
        parts = []
        for p in partitions:
            if isinstance(p, Partition):
                parts.append(p)
            elif isinstance(p, str):
                parts.extend(open_processing_set(p, column))
            else:
                parts.append(memory_partition(p, column))
    if not parts:
        raise ValueError("no partitions")
    sg = StagedGridder(n_uv, cell, parts[0].n_chan, parts[0].n_pol, chan_mode, support,
                       oversampling, fp32, block_rows, psf=dopsf)
    try:
        for p in parts:
            sg.add(p)
    finally:
        sg.close()
    return sg.grid, sg.sumwt


# ------------------------------------------------------------ measurement
def main(argv=None) -> None:
    """Host-resident gridding through the stager against device-resident
    gridding of the same samples: the benchmark generator (64 antennas, 64
    channels, 4 polarisations, 2048^2 cube, FP32) over --times time steps,
    as one in-memory partition.  Prints one JSON line."""
    import argparse
    import time

    from . import synth

    ap = argparse.ArgumentParser(prog="python -m astroviper_amd.staging")
    ap.add_argument("--times", type=int, default=100)
    ap.add_argument("--block-rows", type=int, default=1 << 16)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)
    dev = _dev.require_gpu()
    cfg = synth.CONFIG2
    arr = synth.ArrayConfig(**{**cfg["array"].__dict__, "n_time": args.times})
    obs = synth.make_observation(arr, cfg["sky"])
    vis, w, flag = synth.host_visibilities(obs, cfg["sky"], cfg["flag_frac"], cfg["weight_range"])
    vis = vis.astype(np.complex64)
    w = w.astype(np.float32)
    c = cfg["cell_arcsec"] * synth.ARCSEC
    n_uv, cell = cfg["n_uv"], (-c, c)
    nt, nb, nchan, npol = w.shape
    V = vis.size
    part = array_partition(vis, obs.uvw, w, flag, obs.freq)

    sg = StagedGridder(n_uv, cell, nchan, npol, "cube", fp32=True, block_rows=args.block_rows,
                       device=dev)
    sg.add(part)  # warm-up: pinned and device buffers, plan, kernel loads
    t_staged = []
    for _ in range(args.reps):
        sg.grid.zero_()
        sg.sumwt.zero_()
        sg.blocks = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sg.add(part)
        torch.cuda.synchronize()
        t_staged.append(time.perf_counter() - t0)
    # device-resident: the same samples already in HBM, one plan
    rows = nt * nb
    plan = _dev.GridPlan(rows, nchan, nchan, n_uv, cell, device=dev, n_baseline=nb)
    uvw_d = torch.from_numpy(np.ascontiguousarray(obs.uvw.reshape(rows, 3))).to(dev)
    freq_d = torch.from_numpy(obs.freq).to(dev)
    v_d = torch.from_numpy(vis.reshape(-1, npol)).to(dev)
    w_d = torch.from_numpy(w.reshape(-1, npol)).to(dev)
    f_d = torch.from_numpy(flag.reshape(-1, npol).astype(np.uint8)).to(dev)
    g = torch.empty((nchan, npol) + tuple(n_uv), dtype=torch.complex64, device=dev)
    s = torch.empty((nchan, npol), dtype=torch.float64, device=dev)
    t_dev = []
    for _ in range(args.reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.set_uvw(uvw_d, freq_d)
        plan.grid(v_d, w_d, f_d, grid=g, sumwt=s)
        torch.cuda.synchronize()
        t_dev.append(time.perf_counter() - t0)
    err = float((g - sg.grid).abs().max() / g.abs().max())
    plan.close()
    sg.close()
    host_bytes = vis.nbytes + w.nbytes + flag.nbytes + obs.uvw.nbytes
    ts, td = float(np.median(t_staged)), float(np.median(t_dev[1:]))
    print(json.dumps({
        "workload": f"{arr.n_ant} antennas x {nt} times x {nchan} chan x {npol} pol = {V} vis, "
                    f"{n_uv[0]}x{n_uv[1]} cube, FP32, 10% flagged",
        "staged_host_resident": {"ms": round(ts * 1e3, 2), "M_vis_per_s": round(V / ts / 1e6, 1),
                                 "host_GB_per_s": round(host_bytes / ts / 1e9, 2),
                                 "blocks": sg.blocks, "block_rows": args.block_rows},
        "device_resident": {"ms": round(td * 1e3, 2), "M_vis_per_s": round(V / td / 1e6, 1)},
        "grid_max_rel_diff": err,
    }))


if __name__ == "__main__":
    main()
