"""Batched uneven workload: many independent image planes, each with its own
number of samples, gridded and degridded in one replay of a captured HIP
graph.

A plane is one (channel, polarisation) image.  The samples of all planes
that share a kernel support are concatenated into one flat list whose
``row_chan`` entry names the plane, so one binning pass and one launch per
kernel serve every plane of that support however uneven the counts: a large
plane simply owns more work units than a small one, and the one-wavefront-
per-unit kernels balance them across the chip.  The per-support tile shape
is picked by :func:`autotune`, which times the candidate tiles on the batch
itself and caches the choice.

The reference grids one plane at a time through ``standard_grid_jit``
(core/imaging/imaging_utils/standard_grid.py:246-432) with ``chan_map`` and
``pol_map`` selecting the plane; this module is the batched form of that
call and of ``dgrid_numba`` (standard_degrid.py:388-485).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import torch

from .device import GridPlan, require_gpu

DEFAULT_TILES = ((16, 16), (16, 32), (32, 16), (32, 32), (16, 64))


@dataclass
class PlaneSet:
    """Device-resident samples of the planes that share one support.

    uvw [n, 3] float64; plane [n] int32 (index into freq / the grid planes);
    freq [n_planes] float64; vis [n, 1] complex; weight [n, 1] real;
    flag [n, 1] uint8."""
    support: int
    uvw: torch.Tensor
    plane: torch.Tensor
    freq: torch.Tensor
    vis: torch.Tensor
    weight: torch.Tensor
    flag: torch.Tensor

    @property
    def n_planes(self) -> int:
        return int(self.freq.numel())

    @property
    def n_samples(self) -> int:
        return int(self.uvw.shape[0])


_TUNED: Dict[tuple, Tuple[int, int]] = {}


def _tune_key(ps: PlaneSet, n_uv, oversampling: int) -> tuple:
    dens = round(math.log2(max(1, ps.n_samples) / max(1, ps.n_planes)))
    return (ps.support, int(n_uv[0]), int(n_uv[1]), oversampling, str(ps.vis.dtype),
            ps.n_planes, dens)


def _time_pair(plan: GridPlan, ps: PlaneSet, grid, sumwt, out, reps: int) -> float:
    def once():
        plan.set_uvw(ps.uvw, ps.freq, None, ps.plane)
        plan.grid(ps.vis, ps.weight, ps.flag, grid=grid, sumwt=sumwt)
        plan.degrid(grid, ps.flag, out)

    once()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        once()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def autotune(ps: PlaneSet, n_uv, cell, oversampling: int = 100,
             candidates: Sequence[Tuple[int, int]] = DEFAULT_TILES, reps: int = 3,
             cache: bool = True) -> Tuple[Tuple[int, int], Dict[Tuple[int, int], float]]:
    """Pick the tile shape with the fastest bin+grid+degrid on this plane set.

    Returns (tile, {tile: ms}).  Choices are cached per (support, grid shape,
    dtype, plane count, log2 samples per plane); a cached choice comes back
    with an empty timing table."""
    key = _tune_key(ps, n_uv, oversampling)
    if cache and key in _TUNED:
        return _TUNED[key], {}
    dev = ps.uvw.device
    grid = torch.empty((ps.n_planes, 1, int(n_uv[0]), int(n_uv[1])), dtype=ps.vis.dtype,
                       device=dev)
    sumwt = torch.empty((ps.n_planes, 1), dtype=torch.float64, device=dev)
    out = torch.empty_like(ps.vis)
    times = {}
    for tile in candidates:
        if tile[0] < ps.support or tile[1] < ps.support:
            continue
        plan = GridPlan(ps.n_samples, ps.n_planes, ps.n_planes, n_uv, cell, ps.support,
                        oversampling, tile=tile, device=dev, row_channels=True)
        try:
            times[tuple(tile)] = _time_pair(plan, ps, grid, sumwt, out, reps)
        except NotImplementedError:
            pass
        finally:
            plan.close()
    if not times:
        raise RuntimeError("autotune: no candidate tile is usable")
    best = min(times, key=times.get)
    if cache:
        _TUNED[key] = best
    return best, times


class BatchedGridder:
    """Grid + degrid of several plane sets (one per support) as one step.

    ``step()`` runs the step eagerly; ``capture()`` records it in a HIP
    graph (torch.cuda.CUDAGraph) and ``replay()`` launches the whole batch
    with one call.  Grids [n_planes, 1, n_u, n_v], sums of weights
    [n_planes, 1] float64 and model visibilities [n, 1] are kept per set;
    the model visibilities are degridded from the set's own grid."""

    def __init__(self, sets: List[PlaneSet], n_uv, cell, oversampling: int = 100,
                 tiles: Optional[Dict[int, Tuple[int, int]]] = None, device=None,
                 concurrent: bool = True):
        self.device = require_gpu(device if device is not None else sets[0].uvw.device)
        self.n_uv = (int(n_uv[0]), int(n_uv[1]))
        self.cell = cell
        # schedule: the set with the most kernel work first
        self.sets = sorted(sets, key=lambda s: -s.n_samples * s.support * s.support)
        self.tiles = {}
        self.plans, self.grids, self.sumwts, self.models = [], [], [], []
        for ps in self.sets:
            tile = (tiles or {}).get(ps.support, (16, 16))
            self.tiles[ps.support] = tuple(tile)
            self.plans.append(GridPlan(ps.n_samples, ps.n_planes, ps.n_planes, self.n_uv, cell,
                                       ps.support, oversampling, tile=tile,
                                       device=self.device, row_channels=True))
            self.grids.append(torch.empty((ps.n_planes, 1) + self.n_uv, dtype=ps.vis.dtype,
                                          device=self.device))
            self.sumwts.append(torch.empty((ps.n_planes, 1), dtype=torch.float64,
                                           device=self.device))
            self.models.append(torch.empty_like(ps.vis))
        self.graph: Optional[torch.cuda.CUDAGraph] = None
        self.concurrent = concurrent
        self.streams = [torch.cuda.Stream(device=self.device) for _ in self.sets]

    @property
    def n_samples(self) -> int:
        return sum(ps.n_samples for ps in self.sets)

    def _run_set(self, i: int) -> None:
        ps, plan = self.sets[i], self.plans[i]
        plan.set_uvw(ps.uvw, ps.freq, None, ps.plane)
        plan.grid(ps.vis, ps.weight, ps.flag, grid=self.grids[i], sumwt=self.sumwts[i])
        plan.degrid(self.grids[i], ps.flag, self.models[i])

    def step(self) -> None:
        """One step.  The plane sets share nothing, so each runs on a stream
        of its own (default priority), forked from and joined back to the
        calling stream: under capture the graph holds them as parallel
        branches, and the short phases of one set (scans, small kernels, the
        tail of a launch) overlap the others' work."""
        if not self.concurrent or len(self.sets) == 1:
            for i in range(len(self.sets)):
                self._run_set(i)
            return
        main = torch.cuda.current_stream(self.device)
        for i, st in enumerate(self.streams):
            st.wait_stream(main)
            with torch.cuda.stream(st):
                self._run_set(i)
        for st in self.streams:
            main.wait_stream(st)

    def capture(self) -> torch.cuda.CUDAGraph:
        """Record one step in a HIP graph.  The step is first run eagerly on
        the capture stream so that every kernel's launch attributes are set
        before recording."""
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            self.step()
        torch.cuda.current_stream(self.device).wait_stream(s)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            self.step()
        torch.cuda.synchronize(self.device)
        self.graph = g
        return g

    def replay(self) -> None:
        if self.graph is None:
            raise RuntimeError("capture() first")
        self.graph.replay()

    def close(self) -> None:
        self.graph = None
        for p in self.plans:
            p.close()


def tuned_gridder(sets: List[PlaneSet], n_uv, cell, oversampling: int = 100,
                  candidates: Sequence[Tuple[int, int]] = DEFAULT_TILES, reps: int = 3):
    """BatchedGridder with each support's tile picked by :func:`autotune`.
    Returns (gridder, {support: {tile: ms}})."""
    tiles, report = {}, {}
    for ps in sets:
        tile, times = autotune(ps, n_uv, cell, oversampling, candidates, reps)
        tiles[ps.support] = tile
        report[ps.support] = {f"{t[0]}x{t[1]}": round(ms, 3) for t, ms in times.items()}
    return BatchedGridder(sets, n_uv, cell, oversampling, tiles), report


def _main(argv=None) -> int:
    """Time the batched workload: 256 planes of 512 x 512, sample counts
    log-uniform in [1e4, 2e6] after 0-60% flagging, supports 5/7/9."""
    import argparse
    import json

    from . import synth

    ap = argparse.ArgumentParser(description=_main.__doc__)
    ap.add_argument("--planes", type=int, default=256)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--min-count", type=float, default=1e4)
    ap.add_argument("--max-count", type=float, default=2e6)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--seed", type=int, default=4)
    args = ap.parse_args(argv)
    dev = require_gpu()
    t0 = time.perf_counter()
    sets, cell, info = synth.uneven_plane_sets(args.planes, args.n, (args.min_count, args.max_count),
                                               seed=args.seed, device=dev)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    bg, report = tuned_gridder(sets, (args.n, args.n), cell)
    bg.step()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.reps):
        bg.step()
    e1.record()
    torch.cuda.synchronize()
    eager_ms = e0.elapsed_time(e1) / args.reps
    bg.capture()
    bg.replay()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(args.reps):
        bg.replay()
    e1.record()
    torch.cuda.synchronize()
    graph_ms = e0.elapsed_time(e1) / args.reps
    n = bg.n_samples
    print(json.dumps({
        "workload": f"{args.planes} planes of {args.n}x{args.n}, counts log-uniform "
                    f"[{args.min_count:g}, {args.max_count:g}] after 0-60% flagging, "
                    "supports 5/7/9",
        "samples": n, "unflagged": info["unflagged"],
        "graph_ms": round(graph_ms, 3), "eager_ms": round(eager_ms, 3),
        "M_vis_per_s_graph": round(n / graph_ms / 1e3, 1),
        "M_vis_per_s_eager": round(n / eager_ms / 1e3, 1),
        "tiles": {str(k): list(v) for k, v in bg.tiles.items()},
        "autotune_ms": report, "generate_s": round(gen_s, 1)}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
