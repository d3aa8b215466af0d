"""Image decorrelation analysis: resolution estimate of a single image (src/nanopyx/core/analysis/decorr.pyx:22-430).

Restated on the host in NumPy with the reference's float32 placement. The image is edge-apodised
(src/nanopyx/core/transform/edges.pyx:13-52) and padded to a power-of-two square; the decorrelation curve d(r) is the
correlation between the spectrum and its phase-only (unit-magnitude) version, accumulated over frequency rings up to
r; the curve is recomputed for ``n_g`` high-pass filtered copies (image minus its Gaussian blur, blur widths spaced
geometrically), twice, the second pass on a range refined from the first; the resolution is 2 * pixel_size / kc with
kc the largest cut-off frequency (position of a local maximum) found in the second pass.

Sequential float32 accumulations of the reference (ring sums, cumulative sums) are reproduced in the same order;
its OpenMP reductions (the edge mean, the spectrum norm) are summed in float64 here. Where the reference reads an
uninitialised score (no curve with a positive kc * A), zeros are used.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.ndimage import gaussian_filter

_f32 = np.float32


def _get_circular_mask(w: int, r2: float) -> np.ndarray:
    """1 inside the disc (y - w/2)^2 + (x - w/2)^2 < r2 w^2 / 4 (src/nanopyx/core/generate/mask.pyx:34-50)."""
    radius = float(r2) * w * w / 4
    c = np.arange(w, dtype=np.float64) - w / 2
    dist = c[:, None] ** 2 + c[None, :] ** 2
    return (dist < radius).astype(np.float32)


def _normalize_fft(fft_real: np.ndarray, fft_imag: np.ndarray) -> np.ndarray:
    """Phase-only spectrum (src/nanopyx/core/transform/normalize.pyx:13-29): F / |F|, 0 where |F| = 0."""
    fr = fft_real.astype(np.float64)
    fi = fft_imag.astype(np.float64)
    mag = np.sqrt(fr * fr + fi * fi)
    out = np.zeros((2,) + fft_real.shape, dtype=np.float32)
    nz = mag != 0
    out[0][nz] = (fr[nz] / mag[nz]).astype(np.float32)
    out[1][nz] = (fi[nz] / mag[nz]).astype(np.float32)
    return out


def _apodize_edges(img: np.ndarray) -> np.ndarray:
    """Cosine roll-off of the 20-pixel border towards the mean of the pixels outside the inscribed circle
    (src/nanopyx/core/transform/edges.pyx:13-52)."""
    img = np.asarray(img, dtype=np.float32)
    height, width = img.shape
    offset = 20
    radius = _f32(width / 2 - offset)
    ys = np.arange(height, dtype=np.float64)[:, None]
    xs = np.arange(width, dtype=np.float64)[None, :]
    dist = ((ys - height / 2) ** 2 + (xs - width / 2) ** 2).astype(np.float32)
    outside = dist > radius * radius
    count = int(outside.sum())
    edge_mean = _f32(img[outside].astype(np.float64).sum() / count) if count else _f32(np.nan)
    x0 = np.abs(xs - width / 2).astype(np.float32)
    y0 = np.abs(ys - height / 2).astype(np.float32)
    ex = np.abs(x0.astype(np.float64) - width / 2)
    ey = np.abs(y0.astype(np.float64) - height / 2)
    border = (ex <= offset) | (ey <= offset)
    d = np.minimum(np.broadcast_to(ex, border.shape), np.broadcast_to(ey, border.shape)).astype(np.float32)
    c = ((np.cos(d.astype(np.float64) * math.pi / offset - math.pi) + 1) / 2).astype(np.float32)
    out = img.copy()
    rolled = (c * (img - edge_mean)).astype(np.float32) + edge_mean
    out[border] = rolled[border]
    return out


def _get_max(arr, x1, x2):
    out = [float(x1), arr[x1]]
    for k in range(x1, x2):
        if arr[k] > out[1]:
            out = [float(k), arr[k]]
    return np.asarray(out, dtype=np.float32)


def _get_min(arr, x1, x2):
    out = [float(x1), arr[x1]]
    for k in range(x1, x2):
        if arr[k] < out[1]:
            out = [float(k), arr[k]]
    return np.asarray(out, dtype=np.float32)


class DecorrAnalysis:
    """Same constructor, ``run_analysis`` and result attributes as the reference class (decorr.pyx:22-71, :389-413)."""

    def __init__(self, rmin: float = 0, rmax: float = 1, n_r: int = 50, n_g: int = 10, pixel_size: float = 1,
                 units: str = "pixel", roi: tuple = (0, 0, 0, 0)):
        self.img = None
        self.img_ref = None
        self.rmin = self.rmin2 = _f32(rmin)
        self.rmax = self.rmax2 = _f32(rmax)
        self.n_r = int(n_r)
        self.n_g = int(n_g)
        self.pixel_size = _f32(pixel_size)
        self.units = units
        self.x0, self.x1, self.y0, self.y1 = roi
        self.d0 = np.zeros(self.n_r, dtype=np.float32)
        self.d = np.zeros((self.n_r, 2 * self.n_g), dtype=np.float32)
        self.kc = np.zeros(2 * self.n_g, dtype=np.float32)
        self.a_g = np.zeros(2 * self.n_g, dtype=np.float32)
        self.kc0 = self.a0 = self.kc_gm = self.agm = self.kc_max = self.a_max = _f32(0)
        self.resolution = _f32(0)

    def run_analysis(self, img: np.ndarray):
        self.img = np.asarray(img).astype(np.float32)
        return self._run_analysis()

    # -- pieces (decorr.pyx:73-285) --------------------------------------------------------------------------------
    @staticmethod
    def _get_preprocessed_image(img: np.ndarray) -> np.ndarray:
        """Power-of-two square filled with the first pixel's value; the image occupies rows [0, h) and columns
        [(n - w) // 2, (n - w) // 2 + w) (the reference centres only along x, :73-111)."""
        h, w = img.shape
        n = int(math.pow(2, math.ceil(math.log(max(h, w)) / math.log(2.0))))
        out = np.full((n, n), np.asarray(img).ravel()[0], dtype=np.float32)
        ox = (n - w) // 2
        out[:h, ox:ox + w] = img
        return out

    @staticmethod
    def _linmap(val, valmin, valmax, mapmin, mapmax):
        out = (val - valmin) / _f32(valmax - valmin)
        return (out * _f32(mapmax - mapmin) + _f32(mapmin)).astype(np.float32)

    @staticmethod
    def _get_corr_coef_norm(fft_real, fft_imag, mask) -> np.float32:
        m = mask == 1
        fr = fft_real[m].astype(np.float32)
        fi = fft_imag[m].astype(np.float32)
        c = (fr * fr + fi * fi).astype(np.float64).sum()
        return _f32(math.sqrt(float(_f32(c))))

    @staticmethod
    def _get_best_score(kc, a):
        gm = (kc * a).astype(np.float32)
        out = np.zeros(3, dtype=np.float32)
        gm_max = _f32(0)
        for k in range(kc.shape[0]):
            if gm[k] > gm_max:
                gm_max = gm[k]
                out[:] = (kc[k], a[k], k)
        return out

    @staticmethod
    def _get_max_score(kc, a):
        out = np.zeros(3, dtype=np.float32)
        kc_max = _f32(0)
        for k in range(kc.shape[0]):
            if kc[k] > kc_max:
                kc_max = kc[k]
                out[:] = (kc[k], a[k], k)
        return out

    def _get_corr_coef_ring(self, fft_real, fft_imag, nfft_real, nfft_imag, crmin, crmax) -> np.ndarray:
        """Ring sums of F . Fn and |Fn|^2 over (2n_r) bins (:169-203). The reference walks x outer, y inner, and
        stops at the first position with x*height + y > width*height/2 + height/2; bins are sequential float32
        sums in that order."""
        n_r = self.n_r
        height, width = self.img_ref.shape
        out = np.zeros(2 * n_r, dtype=np.float32)
        crmin, crmax = _f32(crmin), _f32(crmax)
        one_m = _f32(1) - crmax
        ox = int(_f32(width) * one_m / _f32(2))
        oy = int(_f32(height) * one_m / _f32(2))
        w = int(_f32(width) * crmax)
        h = int(_f32(height) * crmax)
        if w <= 0 or h <= 0:
            return out
        xs = np.arange(ox, ox + w)
        ys = np.arange(oy, oy + h)
        X = np.repeat(xs, h)                  # x outer
        Y = np.tile(ys, w)                    # y inner
        k = X * height + Y
        keep = k <= width * height / 2 + height / 2
        X, Y = X[keep], Y[keep]
        dist = ((X - width / 2) ** 2 + (Y - height / 2) ** 2).astype(np.float32)
        dist = np.sqrt((_f32(4) * dist / _f32(width * width)).astype(np.float64)).astype(np.float32)
        inside = (dist >= 0) & (dist <= crmax)
        X, Y, dist = X[inside], Y[inside], dist[inside]
        dist = self._linmap(dist, crmin, crmax, _f32(0), _f32(n_r - 1))
        dist = np.where(dist < 0, _f32(0), dist)
        d = np.floor(dist.astype(np.float64) + 0.5).astype(np.int64)   # C round(): halves away from zero (dist >= 0)
        ok = d < n_r
        X, Y, d = X[ok], Y[ok], d[ok]
        fr, fi = fft_real[Y, X], fft_imag[Y, X]
        nr, ni = nfft_real[Y, X], nfft_imag[Y, X]
        np.add.at(out, d, (fr * nr + fi * ni).astype(np.float32))
        np.add.at(out, d + n_r, (nr * nr + ni * ni).astype(np.float32))
        return out

    def _curve(self, coef, cr) -> np.ndarray:
        """d(k) = sqrt(2) * sum_{n<=k} num / (cr * sqrt(sum_{n<=k} den)); NaN where cr or the sum is 0, and
        d(0) = 0 when NaN (:205-229, :279-292)."""
        n_r = self.n_r
        d = np.cumsum(coef[:n_r], dtype=np.float32)
        c = np.cumsum(coef[n_r:2 * n_r], dtype=np.float32)
        with np.errstate(divide="ignore", invalid="ignore"):
            out = (math.sqrt(2) * d.astype(np.float64) / (float(cr) * np.sqrt(c.astype(np.float64)))).astype(
                np.float32)
        out[(c == 0) | (cr == 0)] = np.nan
        if np.isnan(out[0]):
            out[0] = 0
        return out

    def _compute_d0(self, fft_real, fft_imag) -> np.ndarray:
        nfft = _normalize_fft(fft_real, fft_imag)
        mask = _get_circular_mask(fft_real.shape[1], 1)
        cr = self._get_corr_coef_norm(fft_real, fft_imag, mask)
        coef = self._get_corr_coef_ring(fft_real, fft_imag, nfft[0], nfft[1], self.rmin, self.rmax)
        return self._curve(coef, cr)

    def _get_d_corr_max(self, d, r1, r2) -> np.ndarray:
        """Position and height of the highest local maximum of d, skipping maxima that sit on a monotonic rise at
        the end of the curve (:231-260); the position is mapped onto [r1, r2]."""
        n_r = self.n_r
        dt = _f32(0.001)
        t = np.array(d, dtype=np.float32, copy=True)
        out = _get_max(d, 0, n_r)
        d_length = t.shape[0]
        while out[0] == d_length - 1:
            t[d_length - 1] = 0
            d_length -= 1
            if d_length == 0:
                out[0] = 0
                out[1] = 0
                break
            out = _get_max(t, 0, n_r)
            temp_min = _get_min(t, int(out[0]), d_length - 1)
            if _f32(t[int(out[0])] - temp_min[1]) > dt:
                break
            t[int(out[0])] = temp_min[1]
            out[0] = d_length - 1
        out[0] = _f32(_f32(r1) + _f32(_f32(r2) - _f32(r1)) * out[0] / _f32(n_r - 1))
        return out

    def _compute_d(self) -> np.ndarray:
        """Two passes of n_g high-pass curves; the first narrows the frequency range and the blur widths, the
        second gives kc_gm / kc_max (:262-385)."""
        n_r, n_g = self.n_r, self.n_g
        d_curve = np.zeros((n_r, 2 * n_g), dtype=np.float32)
        g_max = _f32(self.img_ref.shape[1] / 2) if self.kc0 == 0 else _f32(2 / self.kc0)
        g_min = _f32(0.14)
        crmin, crmax = self.rmin, self.rmax
        mask = _get_circular_mask(self.img_ref.shape[1], 1)
        count = 0
        for refine in range(2):
            for k in range(n_g):
                sig = math.exp(math.log(g_min) + (math.log(g_max) - math.log(g_min)) * float(_f32(k) / _f32(n_g - 1)))
                base = np.array(self.img_ref, dtype=np.float32, copy=True)
                blurred = base - gaussian_filter(np.array(self.img_ref, copy=True), sig)
                fft = np.fft.fftshift(np.fft.fft2(blurred)).astype(np.complex64)
                fft_real = np.real(fft).astype(np.float32)
                fft_imag = np.imag(fft).astype(np.float32)
                nfft = _normalize_fft(fft_real, fft_imag)
                cr = self._get_corr_coef_norm(fft_real, fft_imag, mask)
                coef = self._get_corr_coef_ring(fft_real, fft_imag, nfft[0], nfft[1], crmin, crmax)
                d_curve[:, count] = self._curve(coef, cr)
                count += 1
            if refine == 0:
                kc = np.zeros(n_g + 1, dtype=np.float32)
                a = np.zeros(n_g + 1, dtype=np.float32)
                for j in range(n_g):
                    res = self._get_d_corr_max(d_curve[:, j].copy(), crmin, crmax)
                    kc[j] = self.kc[j] = res[0]
                    a[j] = self.a_g[j] = res[1]
                res = self._get_d_corr_max(self.d0, crmin, crmax)
                kc[n_g] = self.kc[n_g] = res[0]
                a[n_g] = self.a_g[n_g] = res[1]
                gm = self._get_best_score(kc, a)
                mx = self._get_max_score(kc, a)
                crmin = _f32(min(float(gm[0]), float(mx[0])) - 0.05)
                if crmin < self.rmin:
                    crmin = self.rmin
                crmax = _f32(max(float(gm[0]), float(mx[0])) + 0.3)
                if crmax > self.rmax:
                    crmax = self.rmax
                self.rmin2, self.rmax2 = crmin, crmax
                ind1 = _f32(min(float(gm[2]), float(mx[2])) - 1)
                ind2 = _f32(max(float(gm[2]), float(mx[2])))
                if ind2 < n_g:
                    lo, hi = math.log(g_min), math.log(g_max)
                    g_temp = math.exp(lo + (hi - lo) * float(ind1 / _f32(n_g - 1)))
                    g_max = _f32(math.exp(lo + (hi - lo) * float(ind2 / _f32(n_g - 1))))
                    g_min = _f32(g_temp)
                else:
                    g_max = g_min
                    g_min = _f32(2 / self.img_ref.shape[1])
            else:
                kc = np.zeros(n_g, dtype=np.float32)
                a = np.zeros(n_g, dtype=np.float32)
                for j in range(n_g):
                    res = self._get_d_corr_max(d_curve[:, j + n_g].copy(), crmin, crmax)
                    kc[j] = self.kc[j] = res[0]
                    a[j] = self.a_g[j] = res[1]
                gm = self._get_best_score(kc, a)
                mx = self._get_max_score(kc, a)
                self.kc_gm, self.agm = gm[0], gm[1]
                self.kc_max, self.a_max = mx[0], mx[1]
        return d_curve

    def _run_analysis(self):
        # the reference crops the ROI into img_ref but then overwrites img_ref with the padded copy of the whole
        # apodised image (:398-405), so the ROI does not change the result
        img_f = _apodize_edges(self.img)
        temp = self._get_preprocessed_image(img_f)
        self.img_ref = temp.copy()
        img_fft = np.fft.fftshift(np.fft.fft2(temp)).astype(np.complex64)
        fft_real = np.real(img_fft).astype(np.float32)
        fft_imag = np.imag(img_fft).astype(np.float32)
        fft_real[fft_real.shape[0] // 2, fft_real.shape[1] // 2] = 0
        fft_imag[fft_imag.shape[0] // 2, fft_imag.shape[1] // 2] = 0
        self.d0 = self._compute_d0(fft_real, fft_imag)
        out = self._get_d_corr_max(self.d0, 0, 1)
        self.kc0, self.a0 = out[0], out[1]
        self.d = self._compute_d()
        with np.errstate(divide="ignore"):
            self.resolution = _f32(2) * self.pixel_size / _f32(self.kc_max)
