"""Staged gridding of processing sets: partitions read from zarr v2 stores
(or memory) stream through pinned buffers in row blocks, with one plan per
stream, and must grid exactly what the oracle grids from the whole arrays
(with the reference's masking, standard_grid.py:660-666: WEIGHT zeroed where
FLAG is set or UVW is NaN).  Small blocks force several blocks per partition,
a partial last block (padded with NaN uvw rows) and slot reuse."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

C = 299792458.0


def rel(a, b):
    den = np.max(np.abs(b))
    return float(np.max(np.abs(np.asarray(a) - b)) / (den if den > 0 else 1.0))


def partitions(n_time=(40, 23, 31), seed=4):
    """Three C1-shaped partitions (27 antennas, 4 channels, 2 pols) at
    different hour angles, 10% flagged, weights U(0.5, 1.5), a few rows of NaN
    uvw; dense cells so that most samples fall on the 256 grid."""
    from astroviper_amd import synth
    out = []
    for k, nt in enumerate(n_time):
        arr = synth.ArrayConfig(seed=seed, n_time=nt, ha_range_h=(-2.0 + k, -1.0 + k))
        obs = synth.make_observation(arr, synth.SkyConfig())
        vis, w, flag = synth.host_visibilities(obs, synth.SkyConfig(), flag_frac=0.1,
                                               weight_range=(0.5, 1.5))
        uvw = obs.uvw.copy()
        uvw[1, 3:9, 0] = np.nan
        out.append(dict(vis=vis, uvw=uvw, w=w, flag=flag, freq=obs.freq))
    maxuv = max(np.nanmax(np.hypot(p["uvw"][..., 0], p["uvw"][..., 1])) for p in out)
    c = 0.45 * C / out[0]["freq"][-1] / maxuv
    return out, (256, 256), (-c, c)


def oracle_grid(oracle, parts, n_uv, cell, mode, psf=False):
    g = sw = None
    for p in parts:
        w = p["w"].copy()
        w[p["flag"]] = 0
        nan = np.isnan(p["uvw"][:, :, 0]) | np.isnan(p["uvw"][:, :, 1])
        w[nan] = 0
        gi, si = oracle.standard_grid(None if psf else p["vis"], p["uvw"], w, p["freq"], n_uv,
                                      cell, chan_mode=mode, do_psf=psf)
        g, sw = (gi, si) if g is None else (g + gi, sw + si)
    return g, sw


@pytest.mark.parametrize("mode", ["continuum", "cube"])
@pytest.mark.parametrize("prec,codec", [(64, None), (32, "zlib")])
def test_staged_processing_set_matches_oracle(gpu, oracle, tmp_path, mode, prec, codec):
    from astroviper_amd import staging as S
    parts, n_uv, cell = partitions()
    if prec == 32:  # the FP32 path grids complex64 / float32 inputs: give the oracle the same
        for p in parts:
            p["vis"] = p["vis"].astype(np.complex64).astype(np.complex128)
            p["w"] = p["w"].astype(np.float32).astype(np.float64)
    for i, p in enumerate(parts):
        S.write_partition(str(tmp_path / "ps" / f"p{i}"), p["vis"], p["uvw"], p["w"], p["flag"],
                          p["freq"], time_chunk=7, compressor=codec)
    gref, swref = oracle_grid(oracle, parts, n_uv, cell, mode)
    nb = parts[0]["w"].shape[1]
    g, sw = S.grid_processing_set(str(tmp_path / "ps"), n_uv, cell, chan_mode=mode,
                                  fp32=prec == 32, block_rows=9 * nb)
    torch.cuda.synchronize()
    e = rel(g.cpu().numpy(), gref)
    print(f"staged {mode} FP{prec}: grid {e:.2e}", flush=True)
    assert e < (1e-4 if prec == 32 else 1e-12)
    np.testing.assert_allclose(sw.cpu().numpy(), swref, rtol=1e-12)
    # the same partitions from memory, and the PSF
    from astroviper_amd import msv4
    dss = [msv4.make_visibility_dataset(p["vis"], p["uvw"], p["w"], p["flag"], p["freq"])
           for p in parts]
    g2, sw2 = S.grid_processing_set(dss, n_uv, cell, chan_mode=mode, fp32=prec == 32,
                                    block_rows=5 * nb)
    assert rel(g2.cpu().numpy(), gref) < (1e-4 if prec == 32 else 1e-12)
    pref, psref = oracle_grid(oracle, parts, n_uv, cell, mode, psf=True)
    pg, psw = S.grid_processing_set(dss, n_uv, cell, chan_mode=mode, dopsf=True,
                                    fp32=prec == 32, block_rows=11 * nb)
    assert rel(pg.cpu().numpy(), pref) < (1e-4 if prec == 32 else 1e-12)
    np.testing.assert_allclose(psw.cpu().numpy(), psref, rtol=1e-12)
