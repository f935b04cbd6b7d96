"""The zarr v2 reader and processing-set discovery of staging.py (CPU only).

The reader is pinned against stores assembled by hand from the zarr v2
specification (a .zarray header and C-order chunk files, edge chunks stored
full size), independently of the module's own writer, and the writer is
checked by reading back through the reader.
"""
import json
import os
import zlib

import numpy as np
import pytest

from astroviper_amd import staging as S


def handmade_store(path, data, chunks, codec=None):
    os.makedirs(path)
    meta = {"zarr_format": 2, "shape": list(data.shape), "chunks": list(chunks),
            "dtype": data.dtype.str, "order": "C", "filters": None, "fill_value": 0,
            "compressor": None if codec is None else {"id": codec, "level": 5}}
    with open(os.path.join(path, ".zarray"), "w") as f:
        json.dump(meta, f)
    grid = [-(-s // c) for s, c in zip(data.shape, chunks)]
    for idx in np.ndindex(*grid):
        blk = np.zeros(chunks, data.dtype)
        sl = tuple(slice(i * c, (i + 1) * c) for i, c in zip(idx, chunks))
        part = data[sl]
        blk[tuple(slice(0, n) for n in part.shape)] = part
        raw = blk.tobytes(order="C")
        if codec == "zlib":
            raw = zlib.compress(raw, 5)
        with open(os.path.join(path, ".".join(map(str, idx))), "wb") as f:
            f.write(raw)


@pytest.mark.parametrize("codec", [None, "zlib"])
def test_reader_matches_handmade_store(tmp_path, codec):
    rng = np.random.default_rng(1)
    data = (rng.standard_normal((11, 7, 5, 2)) + 1j * rng.standard_normal((11, 7, 5, 2)))
    path = str(tmp_path / "vis")
    handmade_store(path, data, (4, 3, 5, 2), codec)
    a = S.ZarrV2Array(path)
    assert a.shape == data.shape and a.dtype == data.dtype
    np.testing.assert_array_equal(a.read(), data)
    for a0, a1 in ((0, 1), (3, 9), (8, 11), (10, 11)):
        np.testing.assert_array_equal(a.read(a0, a1), data[a0:a1])


def test_reader_fills_missing_chunks_and_refuses_blosc(tmp_path):
    data = np.arange(24, dtype=np.float32).reshape(6, 4)
    path = str(tmp_path / "w")
    handmade_store(path, data, (2, 4))
    os.remove(os.path.join(path, "1.0"))  # a chunk never written holds the fill value
    out = S.ZarrV2Array(path).read()
    expect = data.copy()
    expect[2:4] = 0
    np.testing.assert_array_equal(out, expect)
    meta = json.load(open(os.path.join(path, ".zarray")))
    meta["compressor"] = {"id": "blosc", "cname": "lz4", "clevel": 5, "shuffle": 1}
    json.dump(meta, open(os.path.join(path, ".zarray"), "w"))
    with pytest.raises(NotImplementedError, match="blosc"):
        S.ZarrV2Array(path)


@pytest.mark.parametrize("codec", [None, "zlib", "gzip"])
def test_writer_round_trip(tmp_path, codec):
    rng = np.random.default_rng(2)
    for dt, shape, chunks in ((np.complex64, (9, 5, 3, 4), (4, 5, 3, 4)),
                              (np.bool_, (9, 5, 3, 4), (2, 5, 3, 4)),
                              (np.float64, (9, 5, 3), (5, 5, 3))):
        data = rng.standard_normal(shape) > 0 if dt == np.bool_ else \
            rng.standard_normal(shape).astype(dt)
        path = str(tmp_path / f"{np.dtype(dt).name}_{codec}")
        S.write_zarr_v2(path, data, chunks, codec)
        np.testing.assert_array_equal(S.ZarrV2Array(path).read(), data)
        np.testing.assert_array_equal(S.ZarrV2Array(path).read(3, 7), data[3:7])


def test_processing_set_discovery(tmp_path):
    rng = np.random.default_rng(3)
    for name, nt in (("part_b", 5), ("part_a", 3)):
        nb, nf, npol = 4, 2, 2
        S.write_partition(str(tmp_path / "ps" / name / "correlated_xds"),
                          np.ones((nt, nb, nf, npol), np.complex64),
                          rng.standard_normal((nt, nb, 3)), np.ones((nt, nb, nf, npol)),
                          np.zeros((nt, nb, nf, npol), bool), [1e9, 1.1e9], time_chunk=2)
    parts = S.open_processing_set(str(tmp_path / "ps"))
    assert [p.n_time for p in parts] == [3, 5]  # name order
    blk = parts[1].read_times(1, 4)
    assert blk["VISIBILITY"].shape == (3, 4, 2, 2) and blk["UVW"].shape == (3, 4, 3)
    np.testing.assert_array_equal(parts[0].freq, [1e9, 1.1e9])
