"""Fourier interpolation of 2-D images or 3-D stacks on the GPU, with the signatures of
src/nanopyx/core/transform/interpolation.py:189-292 (``fht_space_interpolation`` — the function eSRRF's use_fht
branch calls — and ``fht_space_interpolation_tiled``). Frames larger than the tile size (0 = 256) on either side are
interpolated in Hann-blended tiles, as the reference does. Each call is libnpx_hip ``npx_fht_interpolate``
(batched float64 hipFFT transforms plus gfx950 kernels for the padding, spectrum placement, range rescale and
blending); NumPy inputs return NumPy arrays, device tensors return device tensors.
"""
from __future__ import annotations

import numpy as np

from ... import _hip
from ._le_esrrf import _is_tensor


def _run(image, magnification, mirror, tile_size, overlap):
    a = image
    if not _is_tensor(a):
        a = np.asarray(a)
    if a.ndim not in (2, 3):
        raise ValueError("Input must be 2-D or 3-D array")
    out = _hip.fht_interpolate(a if a.ndim == 3 else a[None], int(magnification), mirror, int(tile_size),
                               int(overlap))
    out = out if a.ndim == 3 else out[0]
    return out if _is_tensor(image) else out.cpu().numpy()


def fht_space_interpolation(image, magnification: int = 2, doMirrorPadding: bool = True, tile_size: int = 0,
                            overlap: int = 16):
    return _run(image, magnification, doMirrorPadding, tile_size, overlap)


def fht_space_interpolation_tiled(image, magnification: int = 2, tile_size: int = 256, overlap: int = 16,
                                  doMirrorPadding: bool = True):
    """Always the tiled path for frames larger than ``tile_size`` (the reference's doMirrorPadding is ignored:
    tiles are always mirror padded); smaller frames are interpolated whole with mirror padding."""
    return _run(image, magnification, True, tile_size, overlap)
