"""Macro-pixel correction on the GPU (src/nanopyx/core/transform/mpcorrector.pyx:7-73): per slice, the mean of
img / (rows*cols) over the macro-pixels of each of the M x M phases, then every pixel divided by its phase
value when that value is non-zero. Operates on a float32 copy, like the reference."""
from __future__ import annotations

import numpy as np

from ... import _hip
from ._le_esrrf import _is_tensor


def macro_pixel_corrector(img, magnification: int):
    import torch
    if _is_tensor(img):
        x = img.to(device="cuda", dtype=torch.float32).clone()
    else:
        x = torch.from_numpy(np.array(img, dtype=np.float32, copy=True)).cuda()
    squeeze = x.dim() == 2
    if squeeze:
        x = x.unsqueeze(0)
    _hip.macro_pixel_(x.contiguous(), int(magnification))
    x = x[0] if squeeze else x
    return x if _is_tensor(img) else x.cpu().numpy()
