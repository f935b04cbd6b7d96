"""Host-side logic that runs without a GPU: the MSv4 container, argument
handling, and that the product refuses to run (instead of falling back to
the CPU) when no GPU is visible."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_visibility_dataset_shape_and_mutation():
    from astroviper_amd.msv4 import is_datatree, make_visibility_dataset
    vis = np.zeros((3, 5, 4, 2), complex)
    ds = make_visibility_dataset(vis, np.zeros((3, 5, 3)), np.ones(vis.shape),
                                 np.zeros(vis.shape, bool), np.linspace(1e9, 1.1e9, 4))
    assert is_datatree(ds) and not is_datatree("x")
    assert ds.dims == {"time": 3, "baseline_id": 5, "frequency": 4, "polarization": 2,
                       "uvw_label": 3}
    assert ds.VISIBILITY.dims == ("time", "baseline_id", "frequency", "polarization")
    w = ds.WEIGHT.data
    w[ds.FLAG.data] = 0.0  # in-place mutation reaches the dataset
    ds.WEIGHT.data[0, 0, 0, 0] = 7.0
    assert ds["WEIGHT"].data[0, 0, 0, 0] == 7.0
    assert "VISIBILITY_MODEL" not in ds
    assert len(ds.coords["frequency"].values) == 4 and ds.frequency.data[0] == 1e9


def test_data_array_selection_views():
    from astroviper_amd.msv4 import DataArray
    da = DataArray(np.zeros((2, 3, 4)), dims=("time", "frequency", "u"),
                   coords={"time": [10.0, 20.0], "frequency": [1, 2, 3]})
    v = da.sel(time=20.0).data
    v[...] = 5.0
    assert np.all(da.data[1] == 5.0) and np.all(da.data[0] == 0.0)
    assert da.isel(frequency=1).dims == ("time", "u")
    with pytest.raises(KeyError):
        da.sel(time=30.0)


def test_product_refuses_to_run_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from astroviper_amd import device
    from astroviper_amd.standard_grid import grid2image_spheroid_ms4
    with pytest.raises(RuntimeError, match="needs a ROCm GPU"):
        device.ps_kernel_1d(100, 7)
    with pytest.raises(RuntimeError, match="needs a ROCm GPU"):
        device.GridPlan(10, 1, 1, (64, 64), (1e-5, 1e-5))
    from astroviper_amd.msv4 import make_visibility_dataset
    vis = np.ones((1, 2, 1, 1), complex)
    ds = make_visibility_dataset(vis, np.zeros((1, 2, 3)), np.ones(vis.shape),
                                 np.zeros(vis.shape, bool), [1e9])
    with pytest.raises(RuntimeError, match="needs a ROCm GPU"):
        grid2image_spheroid_ms4(ds, np.zeros((1, 1, 64, 64)), np.array([-1e-5, 1e-5]))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_synthetic_generator_is_seeded():
    from astroviper_amd import synth
    a = synth.make_observation(synth.ArrayConfig(), synth.SkyConfig())
    b = synth.make_observation(synth.ArrayConfig(), synth.SkyConfig())
    np.testing.assert_array_equal(a.uvw, b.uvw)
    assert a.uvw.shape == (100, 351, 3) and a.freq.shape == (4,)
    assert np.allclose(a.freq[[0, -1]], [1.0e9, 1.4e9])
    va, wa, fa = synth.host_visibilities(a, synth.SkyConfig(), flag_frac=0.1)
    vb, wb, fb = synth.host_visibilities(b, synth.SkyConfig(), flag_frac=0.1)
    np.testing.assert_array_equal(va, vb)
    assert va.shape == (100, 351, 4, 2) and 0.05 < fa.mean() < 0.15
