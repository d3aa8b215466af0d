"""Fourier (zero-padded spectrum) interpolation of 2D images, as the error map uses it.

Restates ``fht_space_interpolation`` of src/nanopyx/core/transform/_interpolation.pyx:201-219 with its helpers
``_mirror_padding_even_square_c`` (:130-170) and ``_fht_space_interpolation_c`` (:173-199): the image is mirror-padded
to a power-of-two square, its centred spectrum is embedded in a magnification-times larger zero spectrum, the
inverse transform's real part is rescaled to the input's value range, and the padding is cropped away. Host NumPy
code: one reference image per parameter sweep.
"""
from __future__ import annotations

import math

import numpy as np
import numpy.fft as fft


def _mirror_padding_even_square(img: np.ndarray) -> np.ndarray:
    h, w = img.shape
    bit_w = int(math.pow(2, math.ceil(math.log2(w))))
    bit_h = int(math.pow(2, math.ceil(math.log2(h))))
    out_dim = max(bit_h, bit_w)
    scale_w = int(math.ceil(out_dim / float(w)))
    if scale_w % 2 == 0:
        scale_w += 1
    scale_h = int(math.ceil(out_dim / float(h)))
    if scale_h % 2 == 0:
        scale_h += 1
    mid_w, mid_h = (scale_w - 1) // 2, (scale_h - 1) // 2
    flip_w = np.arange(w)[::-1]
    flip_h = np.arange(h)[::-1]
    padded = np.zeros((scale_h * h, scale_w * w), dtype=np.float32)
    for sh in range(scale_h):
        rows = np.arange(h) if (sh - mid_h) % 2 == 0 else flip_h
        for sw in range(scale_w):
            cols = np.arange(w) if (sw - mid_w) % 2 == 0 else flip_w
            padded[sh * h:(sh + 1) * h, sw * w:(sw + 1) * w] = img[np.ix_(rows, cols)]
    x_roi = (scale_w * w - out_dim) // 2
    y_roi = (scale_h * h - out_dim) // 2
    return padded[y_roi:y_roi + out_dim, x_roi:x_roi + out_dim]


def _fht_interpolate(img: np.ndarray, factor: int, mirror: bool) -> np.ndarray:
    img = np.ascontiguousarray(img, dtype=np.float32)
    orig_min = np.float32(img.min())
    orig_max = np.float32(img.max())
    work = _mirror_padding_even_square(img) if mirror else img
    F = fft.fftshift(fft.fft2(work).astype(np.complex128))
    h, w = F.shape
    h_int, w_int = h * factor, w * factor
    F_int = np.zeros((h_int, w_int), dtype=np.complex128)
    x_roi, y_roi = (w_int - w) // 2, (h_int - h) // 2
    F_int[y_roi:y_roi + h, x_roi:x_roi + w] = F
    out = np.real(fft.ifft2(fft.ifftshift(F_int))).astype(np.float32)
    out_min = np.float32(out.min())
    out_max = np.float32(out.max())
    if out_max != out_min:
        scale = np.float32((orig_max - orig_min) / (out_max - out_min))
        out = ((out - out_min) * scale + orig_min).astype(np.float32)
    if mirror:
        x_roi = (w_int - factor * img.shape[1]) // 2
        y_roi = (h_int - factor * img.shape[0]) // 2
        return out[y_roi:y_roi + factor * img.shape[0], x_roi:x_roi + factor * img.shape[1]]
    return out


def fht_space_interpolation(image, magnification: int = 2, doMirrorPadding: bool = True) -> np.ndarray:
    """Fourier interpolation of a 2D image or of every frame of a 3D stack by an integer factor."""
    a = np.asarray(image)
    if a.ndim == 2:
        return np.ascontiguousarray(_fht_interpolate(a, int(magnification), doMirrorPadding))
    if a.ndim == 3:
        return np.ascontiguousarray(np.stack([_fht_interpolate(f, int(magnification), doMirrorPadding) for f in a]))
    raise ValueError("Input must be 2D or 3D array")
