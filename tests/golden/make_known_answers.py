"""Writes reference_known_answers.json: the inputs and expected outputs that
the reference's own hot-path tests state as literals.  The reference cannot
be imported here (it needs Python >= 3.11, numba and xarray), so the values
are transcribed from its test files; each entry names the file and lines.

    python tests/golden/make_known_answers.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

ANSWERS = {
    # tests/unit/core/imaging/imaging_utils/test_standard_grid.py:15-41, 84-85, 111
    "standard_grid_point_sources": {
        "nant": 200,
        "freq": 1.4e9,
        "maxUV": 2e3,
        # (row, column, flux): mod_im[sources[0], sources[1]] = sources[2]
        "sources": [[110, 115, 1.0], [87, 92, 2.0]],
        "dirty_thresholds": [[87, 92, 1.8], [110, 115, 0.9]],
        "psf_centre_tol": 0.01,
    },
    # tests/unit/core/imaging/imaging_utils/test_standard_degrid.py:51-66
    "sgrid_basic": {
        "uvw": [1.0, 2.0, 3.0],
        "freq": 1.0e9,
        "scale": [1.0e-4, 1.0e-4, 0.0],
        "offset": [500.0, 500.0, 0.0],
        "sampling": 1,
    },
    # tests/unit/core/imaging/imaging_utils/test_standard_degrid.py:26-49, 150
    "dgrid_constant_grid": {"npix": 200, "freq": 1.4e9, "maxUV": 2e3, "n_uvw": 4096,
                            "grid_value": 2.0, "expected_max_abs": 2.0},
    # tests/unit/core/imaging/test_fft_ifft.py:9-24,
    # tests/unit/core/imaging/imaging_utils/test_standard_image_grid_preparation.py:127-142
    "fft": {"roundtrip_tol": 1e-10, "ones_200_dc": 40000.0, "delta_10_max": 10.0 / 200 / 200},
    # create_prolate_spheroidal_kernel_1D(100, 7) at four indices, evaluated
    # from gcf_prolate_spheroidal.py:87-158 to 7 significant digits
    "kernel_table_support7_ov100": {"0": 0.9999997, "100": 0.5732454, "200": 0.08262342,
                                    "299": 2.866338e-05},
}

if __name__ == "__main__":
    with open(os.path.join(HERE, "reference_known_answers.json"), "w") as f:
        json.dump(ANSWERS, f, indent=1, sort_keys=True)
    print("wrote reference_known_answers.json")
