"""Batched uneven workload on the GPU: every plane of a multi-support batch
against the FP64 oracle gridding that plane alone, HIP-graph replays against
the eager step, and the tile autotuner."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL32 = 1e-4
# two FP32 runs whose records reach a bin in different orders (atomic
# cursors): their tile sums differ by FP32 rounding, which for bins of
# thousands of records reaches ~1e-6 of the grid maximum
FP32_ORDER = 1e-5


def rel(a, b):
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


def small_sets(gpu, dtype=torch.complex64):
    from astroviper_amd import synth
    return synth.uneven_plane_sets(n_planes=12, n_uv=128, counts=(2e2, 2e4), n_pol=2,
                                   seed=11, device=gpu, dtype=dtype)


def per_plane_oracle(oracle, ps, p, n_uv, cell):
    sel = (ps.plane == p).nonzero().flatten()
    n = int(sel.numel())
    uvw = ps.uvw[sel].cpu().numpy().reshape(1, n, 3)
    vis = ps.vis[sel].cpu().numpy().astype(np.complex128).reshape(1, n, 1, 1)
    w = ps.weight[sel].cpu().numpy().astype(np.float64).reshape(1, n, 1, 1)
    fl = ps.flag[sel].cpu().numpy().astype(bool).reshape(1, n, 1, 1)
    w[fl] = 0.0
    f = ps.freq[p:p + 1].cpu().numpy()
    g, sw = oracle.standard_grid(vis, uvw, w, f, n_uv, cell, support=ps.support)
    m = oracle.degrid(uvw, fl, f, g, cell, support=ps.support)
    return g[0, 0], sw[0, 0], m.reshape(-1), sel.cpu().numpy()


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_every_plane_matches_oracle(gpu, oracle, dtype):
    from astroviper_amd.batch import BatchedGridder
    sets, cell, info = small_sets(gpu, dtype)
    assert sorted(ps.support for ps in sets) == [5, 7, 9]
    counts = info["plane_counts"]
    assert max(counts) / min(counts) > 5  # uneven
    bg = BatchedGridder(sets, (128, 128), cell, tiles={5: (16, 16), 7: (16, 32), 9: (32, 16)})
    bg.step()
    torch.cuda.synchronize()
    tol = TOL32 if dtype == torch.complex64 else 1e-10
    for ps, g, sw, m in zip(bg.sets, bg.grids, bg.sumwts, bg.models):
        g = g.cpu().numpy()
        sw = sw.cpu().numpy()
        m = m.cpu().numpy().reshape(-1)
        for p in range(ps.n_planes):
            gref, swref, mref, sel = per_plane_oracle(oracle, ps, p, (128, 128), cell)
            # the GPU grid was gridded from single-precision inputs for
            # complex64; the oracle sees the same rounded values
            assert rel(g[p, 0], gref) < tol, (ps.support, p)
            assert abs(sw[p, 0] - swref) <= 1e-12 * swref
            # degridding the plane's own grid: compare with the oracle
            # degridding the oracle grid
            assert rel(m[sel], mref) < (10 * tol), (ps.support, p)


def test_graph_replays_match_eager(gpu):
    from astroviper_amd.batch import BatchedGridder
    sets, cell, _ = small_sets(gpu)
    bg = BatchedGridder(sets, (128, 128), cell)
    bg.step()
    torch.cuda.synchronize()
    ref_g = [g.clone() for g in bg.grids]
    ref_w = [w.clone() for w in bg.sumwts]
    ref_m = [m.clone() for m in bg.models]
    bg.capture()
    for k in range(3):
        for g in bg.grids:
            g.fill_(float("nan"))  # the replay must rewrite everything
        for m in bg.models:
            m.fill_(float("nan"))
        bg.replay()
        torch.cuda.synchronize()
        # eager work between replays
        x = torch.randn(1 << 16, device=gpu).sum().item()
        assert np.isfinite(x)
        # binning fills each bin through atomic cursors, so the order of the
        # records in a bin (and with it the order of the FP32 tile sums and
        # the FP64 weight sums) can differ from run to run: grids and model
        # visibilities agree to FP32 rounding of sums of thousands of
        # records (up to ~1e-6 of the maximum; FP32_ORDER), sums of weights
        # to FP64
        for g, r in zip(bg.grids, ref_g):
            assert torch.isfinite(g).all(), k
            assert rel(g.cpu().numpy(), r.cpu().numpy()) < FP32_ORDER, k
        for w, r in zip(bg.sumwts, ref_w):
            assert rel(w.cpu().numpy(), r.cpu().numpy()) < 1e-12, k
        for m, r in zip(bg.models, ref_m):
            assert torch.isfinite(m).all(), k
            assert rel(m.cpu().numpy(), r.cpu().numpy()) < FP32_ORDER, k
    bg.close()


def test_autotuner_picks_a_measured_tile(gpu):
    from astroviper_amd import batch
    sets, cell, _ = small_sets(gpu)
    ps = sets[1]
    tile, times = batch.autotune(ps, (128, 128), cell, reps=1)
    assert tile in times and times[tile] == min(times.values())
    again, t2 = batch.autotune(ps, (128, 128), cell, reps=1)
    assert again == tile and t2 == {}  # cached
    bg, report = batch.tuned_gridder(sets, (128, 128), cell, reps=1)
    assert set(bg.tiles) == {5, 7, 9}
    bg.step()
    torch.cuda.synchronize()
    bg.close()


@pytest.mark.parametrize("n_uv", [128, 2304])
def test_row_order_does_not_change_results(gpu, n_uv):
    """Rows of one plane stored together (per-tile counters in LDS) and the
    same rows shuffled across planes (every block mixes planes, most records
    take the global bins) give the same records, grids, sums of weights and
    model visibilities.  At 2304 x 2304 with 16 x 16 tiles the per-tile
    counters and cursors do not fit in LDS and every record takes the
    global path."""
    from astroviper_amd.batch import PlaneSet
    from astroviper_amd.device import GridPlan
    sets, cell, _ = small_sets(gpu)
    ps = max(sets, key=lambda s: s.n_samples)
    g = torch.Generator(device=gpu)
    g.manual_seed(5)
    perm = torch.randperm(ps.n_samples, generator=g, device=gpu)
    sh = PlaneSet(ps.support, ps.uvw[perm].contiguous(), ps.plane[perm].contiguous(), ps.freq,
                  ps.vis[perm].contiguous(), ps.weight[perm].contiguous(),
                  ps.flag[perm].contiguous())
    res = []
    for s in (ps, sh):
        plan = GridPlan(s.n_samples, s.n_planes, s.n_planes, (n_uv, n_uv), cell, s.support, 100,
                        tile=(16, 16), device=gpu, row_channels=True)
        plan.set_uvw(s.uvw, s.freq, None, s.plane)
        st = plan.stats()
        grid, sw = plan.grid(s.vis, s.weight, s.flag)
        model = plan.degrid(grid, s.flag)
        torch.cuda.synchronize()
        res.append((st, grid.cpu().numpy(), sw.cpu().numpy(), model.cpu().numpy()))
        plan.close()
    (st0, g0, w0, m0), (st1, g1, w1, m1) = res
    assert st0 == st1
    assert st0["records"] > st0["grid_valid"] > 0
    assert rel(g1, g0) < FP32_ORDER
    assert rel(w1, w0) < 1e-12
    assert rel(m1, m0[perm.cpu().numpy()]) < FP32_ORDER


def test_concurrent_sets_match_sequential(gpu):
    """The plane sets run on streams of their own; the results are those of
    running them one after another (to FP32 rounding of the grids: record
    order within a bin is not fixed)."""
    from astroviper_amd.batch import BatchedGridder
    sets, cell, _ = small_sets(gpu)
    out = []
    for concurrent in (False, True):
        bg = BatchedGridder(sets, (128, 128), cell, concurrent=concurrent)
        bg.step()
        torch.cuda.synchronize()
        out.append(([g.cpu().numpy() for g in bg.grids], [w.cpu().numpy() for w in bg.sumwts],
                    [m.cpu().numpy() for m in bg.models]))
        bg.close()
    (g0, w0, m0), (g1, w1, m1) = out
    for a, b in zip(g0, g1):
        assert rel(b, a) < FP32_ORDER
    for a, b in zip(w0, w1):
        assert rel(b, a) < 1e-12
    for a, b in zip(m0, m1):
        assert rel(b, a) < FP32_ORDER
