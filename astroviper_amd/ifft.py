"""ifft_uv_to_lm on the GPU (mirror of core/imaging/ifft.py:5-18)."""
from __future__ import annotations

from typing import List

import numpy as np
import torch

from . import device as _dev


def ifft_uv_to_lm(ary: np.ndarray, axes: List[int]) -> np.ndarray:
    """real(fftshift(ifft2(ifftshift(ary)))) over `axes`."""
    dev = _dev.require_gpu()
    a = np.asarray(ary)
    cdt = torch.complex64 if a.dtype in (np.float32, np.complex64) else torch.complex128
    t = torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(cdt)
    ax = [int(x) for x in axes]
    out = torch.fft.fftshift(torch.fft.ifft2(torch.fft.ifftshift(t, dim=ax), dim=ax), dim=ax)
    return out.real.cpu().numpy()
