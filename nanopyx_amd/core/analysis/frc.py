"""Fourier ring correlation and the FIRE number (src/nanopyx/core/analysis/frc.pyx:17-294).

Restated on the host with the reference's own NumPy / SciPy pieces (tukey taper, fft2 + fftshift, savgol_filter)
and its float32 arithmetic: both images are cropped to an even square, tapered, transformed; the ring correlation
at radius r averages bilinear samples of the cross- and auto-spectra at angles 0, 1/r, 2/r, ... < 2*pi
(:121-158); the curve is Savitzky-Golay smoothed (window int(0.0707 * length), order 3), intersected with the 1/7
threshold (:186-231), and the FIRE number is 1 / (first intersection's spatial frequency).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.signal import savgol_filter
from scipy.signal.windows import tukey

f32 = np.float32


def pad_w_zeros_2d(img, height, width):
    """Zero-pads img centred in (height, width) (src/nanopyx/core/transform/padding.py:4-22)."""
    out = np.zeros((height, width), dtype=img.dtype)
    h, w = img.shape
    out[(height - h) // 2:(height - h) // 2 + h, (width - w) // 2:(width - w) // 2 + w] = img
    return out


def make_even_square(img):
    """Crops to the largest even square (src/nanopyx/core/analysis/ccm_helper_functions.pyx:36-80)."""
    h, w = img.shape
    if h == w and w % 2 == 0:
        return img
    m = min(w, h)
    if m % 2:
        m -= 1
    h0 = (h - m) // 2
    h1 = h - (h - m) // 2 - 1 if (h - m) % 2 else h - (h - m) // 2
    w0 = int((w - m) / 2)
    w1 = w - (w - m) // 2 - 1 if (w - m) % 2 else w - (w - m) // 2
    return img[h0:h1, w0:w1]


class FIRECalculator:
    def __init__(self, pixel_size: float = 1, units: str = "pixel"):
        self.pixel_size = pixel_size
        self.units = units
        self.threshold = 1 / 7
        self.fire_number = 0.0
        self.field_of_view = 0
        self.frc_curve = None
        self.threshold_curve = None
        self.intersections = None

    @staticmethod
    def _squared_tapered(img):
        tx = tukey(img.shape[1], alpha=0.25).astype(np.float32)
        ty = tukey(img.shape[0], alpha=0.25).astype(np.float32)
        return ((img * tx[None, :]) * ty[:, None]).astype(np.float32)

    @staticmethod
    def calculate_fft(img):
        return np.fft.fftshift(np.fft.fft2(img)).astype(np.complex64)

    def _frc_value(self, centre, size, images):
        max_r = centre - 1
        spatial_frequency = np.linspace(0, 1 / (2 * self.pixel_size), max_r, dtype=np.float32)
        limit = f32(math.pi * 2)
        results = np.zeros((max_r, 3), dtype=np.float32)
        results[0] = (0, 1, 1)
        img0, img1, img2 = images
        for r in range(1, max_r):
            angle_step = f32(1.0 / r)
            angles = np.arange(0, float(limit), float(angle_step), dtype=np.float32)
            cos_a = np.cos(angles.astype(np.float64)).astype(np.float32)
            sin_a = np.sqrt((f32(1) - cos_a * cos_a).astype(np.float64)).astype(np.float32)
            sin_a = np.where(angles.astype(np.float64) > math.pi, -sin_a, sin_a).astype(np.float32)
            rf = f32(r)
            x = (f32(max_r + 1) + rf * cos_a).astype(np.float32)
            y = (f32(max_r + 1) + rf * sin_a).astype(np.float32)
            xb = x.astype(np.int32)  # C truncation
            yb = y.astype(np.int32)
            xf = np.maximum(x - xb.astype(np.float32), f32(0)).astype(np.float32)
            yf = np.maximum(y - yb.astype(np.float32), f32(0)).astype(np.float32)
            ll = yb * size + xb
            sums = []
            for im in (img0, img1, img2):
                lo_l, lo_r = im[ll], im[ll + 1]
                up_l, up_r = im[ll + size], im[ll + size + 1]
                up = (up_l + xf * (up_r - up_l)).astype(np.float32)
                lo = (lo_l + xf * (lo_r - lo_l)).astype(np.float32)
                val = (lo + yf * (up - lo)).astype(np.float32)
                s = f32(0)
                for v in val:  # sequential float32 sum, as the reference's loop
                    s = f32(s + v)
                sums.append(s)
            results[r, 0] = spatial_frequency[r]
            results[r, 1] = f32(float(sums[0]) / math.sqrt(float(f32(sums[1] * sums[2]))))
            results[r, 2] = angles.shape[0]
        return results

    def _calculate_frc_curve(self, img1, img2):
        a = np.array(img1, dtype=np.float32)
        b = np.array(img2, dtype=np.float32)
        h, w = a.shape
        a = pad_w_zeros_2d(a, h, w)
        b = pad_w_zeros_2d(b, h, w)
        if not (a.shape[0] == a.shape[1] and a.shape[1] % 2 == 0):
            a = make_even_square(a)
            b = make_even_square(b)
        a = self._squared_tapered(a)
        b = self._squared_tapered(b)
        f1 = self.calculate_fft(a)
        f2 = self.calculate_fft(b)
        size = f1.shape[0]
        self.field_of_view = size
        a1 = np.real(f1).ravel().astype(np.float32)
        b1 = np.imag(f1).ravel().astype(np.float32)
        a2 = np.real(f2).ravel().astype(np.float32)
        b2 = np.imag(f2).ravel().astype(np.float32)
        images = ((a1 * a2 + b1 * b2).astype(np.float32), (a1 * a1 + b1 * b1).astype(np.float32),
                  (a2 * a2 + b2 * b2).astype(np.float32))
        self.frc_curve = self._frc_value(size // 2, size, images)

    def _smooth(self):
        window = int(0.0707 * self.frc_curve.shape[0])
        if 3 >= window:
            window = 4
        self.frc_curve[:, 1] = savgol_filter(self.frc_curve[:, 1], window_length=window, polyorder=3)

    def _intersections(self):
        frc = self.frc_curve.copy()
        thr = np.full(frc.shape[0], 1.0 / 7.0, dtype=np.float32)
        self.threshold_curve = thr
        out = []
        for i in range(1, frc.shape[0]):
            y1, y2, y3, y4 = frc[i - 1, 1], frc[i, 1], thr[i - 1], thr[i]
            if not ((y3 >= y1 and y4 < y2) or (y1 >= y3 and y2 < y4)):
                continue
            x1, x2 = frc[i - 1, 0], frc[i, 0]
            x3, x4 = x1, x2
            x1_x2, x3_x4 = f32(x1 - x2), f32(x3 - x4)
            y1_y2, y3_y4 = f32(y1 - y2), f32(y3 - y4)
            den = f32(x1_x2 * y3_y4 - y1_y2 * x3_x4)
            if den == 0:
                if y1 == y3:
                    out.append((x1, y1))
            else:
                px = f32(f32(f32(f32(x1 * y2) - f32(y1 * x2)) * x3_x4 - f32(x1_x2 * f32(f32(x3 * y4) - f32(y3 * x4))))
                         / den)
                if x1 <= px < x2:
                    m = f32(f32(y4 - y3) / f32(x2 - x1))
                    py = f32(f32(m * px) + f32(y3 - f32(m * x1)))
                    out.append((px, py))
                else:
                    out.append((px, y1))
        self.intersections = np.array(out, dtype=np.float32).reshape(-1, 2)

    def calculate_fire_number(self, img_1, img_2) -> float:
        self._calculate_frc_curve(np.asarray(img_1).astype(np.float32), np.asarray(img_2).astype(np.float32))
        self._smooth()
        self._intersections()
        self.fire_number = float(np.float32(1 / float(self.intersections[0, 0]))) if self.intersections.shape[0] else 0.0
        return self.fire_number
