"""fft_lm_to_uv on the GPU (mirror of core/imaging/fft.py:5-18)."""
from __future__ import annotations

from typing import List

import numpy as np
import torch

from . import device as _dev


def fft_lm_to_uv(ary: np.ndarray, axes: List[int]) -> np.ndarray:
    """fftshift(fft2(ifftshift(ary))) over `axes`; complex result."""
    dev = _dev.require_gpu()
    a = np.asarray(ary)
    cdt = torch.complex64 if a.dtype in (np.float32, np.complex64) else torch.complex128
    t = torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(cdt)
    ax = [int(x) for x in axes]
    out = torch.fft.fftshift(torch.fft.fft2(torch.fft.ifftshift(t, dim=ax), dim=ax), dim=ax)
    return out.cpu().numpy()
