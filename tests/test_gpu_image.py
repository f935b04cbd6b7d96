"""The dirty-image step of grid2image_spheroid_ms4 (standard_grid.py:693-708):
real(fftshift(ifft2(ifftshift(grid)))) / taper * nx * ny / sumwt, as
device.dirty_image computes it (a complex-to-real FFT of the folded half
spectrum for even sizes, the complex FFT otherwise), against the same
expression in torch FP64 on random, non-Hermitian grids."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def reference(grid, sumwt):
    from astroviper_amd.device import ps_correction_1d
    g = grid.to(torch.complex128)
    n_u, n_v = g.shape[-2:]
    y = torch.fft.fftshift(torch.fft.ifft2(torch.fft.ifftshift(g, dim=(-2, -1)), dim=(-2, -1)),
                           dim=(-2, -1)).real
    taper = ps_correction_1d(n_u, g.device)[:, None] * ps_correction_1d(n_v, g.device)[None, :]
    return y / taper * (n_u * n_v) / sumwt.to(torch.float64)[..., None, None]


@pytest.mark.parametrize("shape", [(2, 3, 96, 128), (1, 2, 128, 128), (2, 1, 95, 64),
                                   (1, 1, 64, 33)])
@pytest.mark.parametrize("cdt", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("out_dtype", [torch.float64, torch.float32])
def test_dirty_image_matches_torch_fp64(gpu, shape, cdt, out_dtype):
    from astroviper_amd.device import dirty_image
    gen = torch.Generator(device="cpu").manual_seed(sum(shape))
    g = torch.complex(torch.randn(shape, generator=gen, dtype=torch.float64),
                      torch.randn(shape, generator=gen, dtype=torch.float64)).to(cdt).to(gpu)
    sw = (torch.rand(shape[:2], generator=gen, dtype=torch.float64) + 0.5).to(gpu)
    img = dirty_image(g, sw, out_dtype=out_dtype)
    ref = reference(g, sw)
    assert img.dtype == out_dtype and img.shape == ref.shape
    e = float((img.to(torch.float64) - ref).abs().max() / ref.abs().max())
    print(f"dirty image {shape} {cdt} -> {out_dtype}: {e:.2e}", flush=True)
    assert e < (1e-13 if out_dtype == torch.float64 else 1e-6)
