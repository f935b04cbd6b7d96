"""Precision sweep at a small size: every variant's errors against the FP64
run meet the FP32 bars, and compensated accumulation is no worse."""
import json

import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_precision_sweep_small(capsys):
    from astroviper_amd import sweep
    rc = sweep.main(["--n", "256", "--n-time", "40", "--chan", "4", "--ant", "16",
                     "--cell-arcsec", "8", "--reps", "1"])
    assert rc == 0
    out = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    v = out["variants"]
    assert set(v) == {"fp64", "fp32", "fp32_compensated"}
    for name in ("fp32", "fp32_compensated"):
        r = v[name]
        assert r["grid_max_rel"] < 1e-4 and r["grid_rms_rel"] < 1e-5
        assert r["sumwt_rel"] < 1e-12
        assert r["degrid_max_rel"] < 1e-4
        assert r["dirty_max_rel_taper_ge_1e-2"] < 1e-4 and r["dirty_max_rel"] < 3e-3
        assert r["psf_max_rel_taper_ge_1e-2"] < 1e-4 and r["psf_max_rel"] < 3e-3
    assert v["fp32_compensated"]["grid_rms_rel"] <= 1.2 * v["fp32"]["grid_rms_rel"]
