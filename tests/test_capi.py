"""libavgrid.so loads on the CPU host and exports every function that
include/avgrid.h declares (no compute calls: no GPU here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "avgrid.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(av_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_abi():
    names = declared_functions()
    for n in ("av_grid", "av_degrid", "av_plan_create", "av_plan_set_uvw", "av_normalise",
              "av_ps_kernel_1d", "av_last_error", "av_grid_plain"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from astroviper_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the whole header
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_error_path_without_gpu_work():
    from astroviper_amd import _lib
    L = _lib.lib()
    rc = L.av_plan_create(10, 1, 1, 64, 64, 1e-5, 1e-5, 99, 100, 0, 0, 0, 0,
                          ctypes.byref(ctypes.c_void_p()))
    assert rc == _lib.AV_ERR_UNSUPPORTED
    assert b"support" in L.av_last_error()
    rc = L.av_plan_create(10, 1, 1, 64, 64, 1e-5, 1e-5, 7, 100, 0, 0, 0, 0, None)
    assert rc == _lib.AV_ERR_INVALID


def test_abi_version_matches_header():
    from astroviper_amd import _lib
    src = open(os.path.join(ROOT, "include", "avgrid.h")).read()
    m = re.search(r"#define AV_ABI_VERSION (\d+)", src)
    assert int(m.group(1)) == _lib.AV_ABI_VERSION == _lib.lib().av_version()
