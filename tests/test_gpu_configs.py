"""GPU parity on subsets of the two large workloads, through the exact kernel
instantiations the benchmark and the precision sweep time.

* 64 antennas (2016 baselines), 4 polarisations, 1" cells, 2048 x 2048 grid,
  16 x 16 tiles, 10% flagged, weights U(0.5, 1.5): the benchmark's generator
  (synth.CONFIG2) over 100 of its 1000 time samples and 2 of its 64 channels
  (1.0 and 1.4 GHz), cube and continuum, FP32 and FP64.  These plans run
  the split-FP16 kernel k_grid_h3<4> (the benchmark's FP32 gridder), the
  FP32-products kernel k_grid_mfma<float,4,1,1>, the sub-tile kernel,
  k_grid_mfma<double,4,1,1>, and the v-half degridders (4 polarisations,
  12 column blocks, 8 per half): split-FP16 k_degrid_h3<4,12,8> (the
  benchmark's), and k_degrid_mfma for FP32 products and FP64.
* The same generator at 4096 x 4096 (the precision sweep's grid), 40 time
  samples x 2 channels, cube: FP32, FP32 compensated and FP64.

Every output the gridder produces is compared with the FP64 oracle (a
loop-by-loop restatement of standard_grid_jit, standard_grid.py:246-432, and
dgrid_numba, standard_degrid.py:388-485) on the same samples: the grid, the
sum of weights, the dirty image and the PSF (standard_grid.py:693-708), and
the degridded model.  The samples are generated as float32/complex64 values
so the FP32 and FP64 runs grid identical inputs and share one oracle run.

Bars (max |x - ref| / max |ref| per output): FP64 grids, sums of weights and
models 1e-12, FP64 images 1e-9 (the taper correction magnifies FFT rounding
up to 6e4 times at the image corners); FP32 grids, models and images 1e-4,
the images over the full field (on these dense grids the FP32 rounding
errors average out; the largest measured is 2e-5), and 1e-6 where the taper
is at least 1e-2.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL64 = 1e-12
IMG64 = 1e-9
TOL32 = 1e-4
INNER32 = 1e-6


def rel(a, b):
    den = np.max(np.abs(b))
    return float(np.max(np.abs(np.asarray(a) - b)) / (den if den > 0 else 1.0))


def image_errors(img, ref):
    """(full field, taper >= 1e-2) max errors relative to max |ref|."""
    from oracle import oracle as O
    _, corr = O.create_prolate_spheroidal_kernel(100, 7, np.array(img.shape[-2:]))
    inner = corr >= 1e-2
    den = np.max(np.abs(ref))
    d = np.abs(img - ref)
    return float(d.max() / den), float(d[..., inner].max() / den)


def make_subset(n_time, chans, n):
    """The benchmark generator restricted to `n_time` time samples (over the
    full hour-angle range) and the channels `chans` of its 64."""
    from astroviper_amd import synth
    cfg = synth.CONFIG2
    arr = synth.ArrayConfig(**{**cfg["array"].__dict__, "n_time": n_time})
    obs = synth.make_observation(arr, cfg["sky"])
    obs.freq = np.ascontiguousarray(obs.freq[list(chans)])
    vis, w, flag = synth.host_visibilities(obs, cfg["sky"], cfg["flag_frac"],
                                           cfg["weight_range"])
    vis = vis.astype(np.complex64).astype(np.complex128)
    w = w.astype(np.float32).astype(np.float64)
    c = cfg["cell_arcsec"] * synth.ARCSEC
    return dict(uvw=obs.uvw, freq=obs.freq, vis=vis, w=w, flag=flag, n_uv=(n, n),
                cell=(-c, c))


def oracle_outputs(oracle, d, mode):
    w0 = d["w"].copy()
    w0[d["flag"]] = 0
    g, sw = oracle.standard_grid(d["vis"], d["uvw"], w0, d["freq"], d["n_uv"], d["cell"],
                                 chan_mode=mode)
    pg, psw = oracle.standard_grid(None, d["uvw"], w0, d["freq"], d["n_uv"], d["cell"],
                                   chan_mode=mode, do_psf=True)
    out = dict(grid=g, sumwt=sw, dirty=oracle.normalise(g, sw), psf=oracle.normalise(pg, psw),
               psf_sumwt=psw)
    return out


def run_gpu(gpu, d, mode, prec, compensated=False, tile=(16, 16), subtile=False,
            f32_products=False, row_order=True):
    from astroviper_amd.device import (GridPlan, degrid_kernel_name, dirty_image,
                                       grid_kernel_name)
    nt, nb, nchan, npol = d["w"].shape
    rows = nt * nb
    cube = mode == "cube"
    # FP32 without compensated sums runs the sub-tile kernel (4x4x1 products)
    # or the FP32-products kernel when asked for, else the split-FP16 kernel;
    # FP64 and compensated FP32 the dense 16x16x4 kernel
    plain32 = prec == 32 and not compensated
    expect = ("k_grid_sub" if plain32 and subtile else
              "k_grid_mfma" if not plain32 or f32_products else "k_grid_h3")
    assert grid_kernel_name(tile, npol, fp64=prec == 64, compensated=compensated,
                            subtile=subtile, f32_products=f32_products) == expect
    assert degrid_kernel_name(tile, npol, 7, fp64=prec == 64, f32_products=f32_products) == (
        "k_degrid_h3" if prec == 32 and not f32_products else "k_degrid_mfma")
    # the benchmark bins its time-major rows baseline by baseline
    plan = GridPlan(rows, nchan, nchan if cube else 1, d["n_uv"], d["cell"], 7, 100,
                    tile=tile, device=gpu, n_baseline=nb if row_order else 1)
    plan.set_uvw(torch.from_numpy(np.ascontiguousarray(d["uvw"].reshape(rows, 3))).to(gpu),
                 torch.from_numpy(d["freq"]).to(gpu))
    rdt = torch.float32 if prec == 32 else torch.float64
    cdt = torch.complex64 if prec == 32 else torch.complex128
    v = torch.from_numpy(d["vis"].reshape(rows * nchan, npol)).to(gpu, cdt)
    w = torch.from_numpy(d["w"].reshape(rows * nchan, npol)).to(gpu, rdt)
    fl = torch.from_numpy(d["flag"].reshape(rows * nchan, npol).astype(np.uint8)).to(gpu)
    g, sw = plan.grid(v, w, fl, compensated=compensated, subtile=subtile,
                      f32_products=f32_products)
    dirty = dirty_image(g, sw)
    pg, psw = plan.grid(None, w, fl, psf=True, compensated=compensated, subtile=subtile,
                        f32_products=f32_products)
    psf = dirty_image(pg, psw)
    torch.cuda.synchronize()
    st = plan.stats()
    out = dict(grid=g.cpu().numpy(), sumwt=sw.cpu().numpy(), dirty=dirty.cpu().numpy(),
               psf=psf.cpu().numpy(), psf_sumwt=psw.cpu().numpy(), plan=plan, flag=fl,
               stats=st)
    return out


def check_outputs(got, ref, prec, label):
    e_g = rel(got["grid"], ref["grid"])
    e_s = rel(got["sumwt"], ref["sumwt"])
    e_ps = rel(got["psf_sumwt"], ref["psf_sumwt"])
    d_full, d_in = image_errors(got["dirty"], ref["dirty"])
    p_full, p_in = image_errors(got["psf"], ref["psf"])
    print(f"{label}: grid {e_g:.2e} sumwt {e_s:.2e} dirty {d_full:.2e} (taper>=1e-2 {d_in:.2e}) "
          f"psf {p_full:.2e} (taper>=1e-2 {p_in:.2e})", flush=True)
    assert e_s < 1e-12 and e_ps < 1e-12, label
    if prec == 64:
        assert e_g < TOL64, label
        assert d_full < IMG64 and p_full < IMG64, label
    else:
        assert e_g < TOL32, label
        assert d_full < TOL32 and p_full < TOL32, label
        assert d_in < INNER32 and p_in < INNER32, label


# ----------------------------------------------------- 2048^2, 4 pols
@pytest.fixture(scope="module")
def c2_subset(oracle):
    d = make_subset(100, (0, 63), 2048)
    refs = {m: oracle_outputs(oracle, d, m) for m in ("cube", "continuum")}
    return d, refs


@pytest.mark.parametrize("mode", ["cube", "continuum"])
@pytest.mark.parametrize("prec,variant", [(32, "split16"), (32, "f32_products"), (32, "subtile"),
                                          (64, "")])
def test_benchmark_workload_subset(gpu, oracle, c2_subset, mode, prec, variant):
    d, refs = c2_subset
    ref = refs[mode]
    got = run_gpu(gpu, d, mode, prec, subtile=variant == "subtile",
                  f32_products=variant == "f32_products")
    st = got["stats"]
    assert st["grid_valid"] > 0.6 * st["samples"]
    check_outputs(got, ref, prec, f"2048^2 {mode} FP{prec} {variant}")
    # degrid the oracle's grid (rounded to the run's precision) with the
    # benchmark's plan, against the oracle's degridder on that same grid
    cdt = torch.complex64 if prec == 32 else torch.complex128
    gin = ref["grid"].astype(np.complex64).astype(np.complex128) if prec == 32 else ref["grid"]
    mref = oracle.degrid(d["uvw"], d["flag"], d["freq"], gin, d["cell"])
    m = got["plan"].degrid(torch.from_numpy(gin).to(gpu, cdt), got["flag"],
                           f32_products=variant == "f32_products").cpu().numpy()
    m = m.reshape(mref.shape)
    e_m = rel(m, mref)
    print(f"2048^2 {mode} FP{prec}: degrid {e_m:.2e}", flush=True)
    assert e_m < (TOL32 if prec == 32 else TOL64)
    assert np.all(m[d["flag"]] == 0)
    got["plan"].close()


@pytest.mark.parametrize("mode", ["cube", "continuum"])
def test_row_order_binning(gpu, oracle, c2_subset, mode):
    """Binning baseline by baseline (av_plan_set_row_order) against row
    order: the same counts, bins and records, so FP64 grids, sums of weights
    and degridded models that differ only by the order of their sums (a
    missing or doubled record would show as an O(1) difference)."""
    d, refs = c2_subset
    a = run_gpu(gpu, d, mode, 64, row_order=True)
    b = run_gpu(gpu, d, mode, 64, row_order=False)
    assert a["stats"] == b["stats"]
    assert rel(a["grid"], b["grid"]) < 1e-14
    assert rel(a["sumwt"], b["sumwt"]) < 1e-14
    check_outputs(a, refs[mode], 64, f"2048^2 {mode} FP64 baseline order")
    gin = torch.from_numpy(refs[mode]["grid"]).to(gpu)
    ma = a["plan"].degrid(gin, a["flag"]).cpu().numpy()
    mb = b["plan"].degrid(gin, b["flag"]).cpu().numpy()
    assert rel(ma, mb) < 1e-14
    a["plan"].close()
    b["plan"].close()


def test_row_order_hint_that_does_not_match(gpu, oracle, c2_subset):
    """A baseline count that divides the rows but is not the rows' real one
    (1600 for 2016 baselines) only changes the order in which a chunk's rows
    are visited: every row is still visited once, so the counts and records
    are those of row order and the FP64 grid agrees to rounding."""
    d, refs = c2_subset
    from astroviper_amd.device import GridPlan
    nt, nb, nchan, npol = d["w"].shape
    rows = nt * nb
    assert nb != 1600 and rows % 1600 == 0
    uvw = torch.from_numpy(np.ascontiguousarray(d["uvw"].reshape(rows, 3))).to(gpu)
    freq = torch.from_numpy(d["freq"]).to(gpu)
    v = torch.from_numpy(d["vis"].reshape(rows * nchan, npol)).to(gpu)
    w = torch.from_numpy(d["w"].reshape(rows * nchan, npol)).to(gpu)
    fl = torch.from_numpy(d["flag"].reshape(rows * nchan, npol).astype(np.uint8)).to(gpu)
    out = []
    for hint in (1600, 1):
        plan = GridPlan(rows, nchan, nchan, d["n_uv"], d["cell"], 7, 100, tile=(16, 16),
                        device=gpu, n_baseline=hint)
        plan.set_uvw(uvw, freq)
        g, sw = plan.grid(v, w, fl)
        torch.cuda.synchronize()
        out.append((plan.stats(), g.cpu().numpy(), sw.cpu().numpy()))
        plan.close()
    (s0, g0, w0), (s1, g1, w1) = out
    assert s0 == s1
    assert rel(g0, g1) < 1e-14 and rel(w0, w1) < 1e-14
    assert rel(g0, refs["cube"]["grid"]) < TOL64


def test_row_order_arguments(gpu):
    """av_plan_set_row_order takes n_baseline >= 1 dividing n_rows; other
    values are refused with an error and leave the plan usable."""
    import ctypes
    from astroviper_amd import _lib
    from astroviper_amd.device import GridPlan
    plan = GridPlan(6000, 2, 2, (256, 256), (-1e-5, 1e-5), 7, 100, tile=(16, 16), device=gpu)
    lib = _lib.lib()
    for bad in (0, -3, 7):  # 7 does not divide 6000
        assert lib.av_plan_set_row_order(plan._h, bad) != 0
        assert b"n_baseline" in lib.av_last_error()
    assert lib.av_plan_set_row_order(plan._h, 300) == 0
    assert lib.av_plan_set_row_order(ctypes.c_void_p(), 300) != 0
    plan.close()


# ------------------------------------------- 4096^2 precision variants
@pytest.fixture(scope="module")
def c5_subset(oracle):
    d = make_subset(40, (0, 63), 4096)
    return d, oracle_outputs(oracle, d, "cube")


@pytest.mark.parametrize("variant", ["fp32", "fp32_products", "fp32_compensated", "fp64"])
def test_precision_sweep_workload_subset(gpu, oracle, c5_subset, variant):
    d, ref = c5_subset
    prec = 64 if variant == "fp64" else 32
    got = run_gpu(gpu, d, "cube", prec, compensated=variant == "fp32_compensated",
                  f32_products=variant == "fp32_products")
    check_outputs(got, ref, prec, f"4096^2 cube {variant}")
    got["plan"].close()
