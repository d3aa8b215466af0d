"""Input normalisation of the eSRRF operators (src/nanopyx/core/transform/__interpolation_tools__.py:4-27)."""
from __future__ import annotations

import numpy as np


def check_image(image, keep_uint16: bool = False):
    """Same checks and errors as the reference's ``check_image``: ndarray, 2D -> (1, H, W), 3D kept,
    other dtypes cast to float32. With ``keep_uint16`` a uint16 stack is passed through unchanged, because the
    gfx950 kernels convert uint16 to float32 in-kernel with the identical result. Device tensors (torch) are
    passed through after the same shape check."""
    try:
        import torch
        if isinstance(image, torch.Tensor):
            if image.dim() not in (2, 3):
                raise ValueError("Image must be 2D and 3D (sequence of 2D images)")
            return image.unsqueeze(0) if image.dim() == 2 else image
    except ImportError:  # pragma: no cover - torch is part of this stack
        pass
    image = np.asarray(image)
    if type(image) is not np.ndarray:
        raise TypeError("Image must be of type np.ndarray")
    if image.ndim != 2 and image.ndim != 3:
        raise ValueError("Image must be 2D and 3D (sequence of 2D images)")
    if image.dtype != np.float32 and not (keep_uint16 and image.dtype == np.uint16):
        image = image.astype(np.float32, copy=False)
    if image.ndim == 2:
        image = image.reshape((1, image.shape[0], image.shape[1]))
    return image
