"""astroviper_amd -- MI355X-native standard gridder with the reference's
entry points (casangi/astroviper core/imaging/imaging_utils).

Host-array API (reference signatures):
    standard_grid.grid2image_spheroid_ms4, standard_grid_numpy_wrap[_input_checked],
    standard_degrid.degrid_spheroid_ms4, make_point_spread_function.make_psf,
    gcf_prolate_spheroidal.*, fft.fft_lm_to_uv, ifft.ifft_uv_to_lm,
    standard_image_grid_preparation.grid2xradio_spheroid_ms4, imager.*
Device API (torch tensors resident on the GPU):
    device.GridPlan, device.dirty_image, device.grid_plain, device.degrid_plain
"""
__version__ = "0.1.0"
