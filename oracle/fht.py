"""ORACLE (test infrastructure only) — NumPy restatement of the Fourier interpolation that eSRRF's use_fht=True branch
calls (``fht_space_interpolation`` of src/nanopyx/core/transform/interpolation.py:189-292, imported as
``dht_interpolation`` by src/nanopyx/core/transform/_le_esrrf.pyx:18), and of that branch (:183-188).

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may import this module.

* ``_mirror_pad`` follows interpolation.py:22-55: mirrored copies of the image tiled to a power-of-two square.
* ``fht_2d`` follows :58-109: float64 rfft2 of the padded image, its quadrants placed in the corners of an
  M-times larger half spectrum, irfft2 to the big grid, float32; rescaled (float32 arithmetic) to the input's range
  when the output is not constant; the padding cropped away.
* ``fht_2d_tiled`` follows :112-186: tiles of ``tile_size`` with ``overlap`` on each side (step tile - 2*overlap;
  one extra tile flush with the end when the last one stops short), each interpolated with mirror padding, blended
  with a float32 Hann window and normalised by the summed window.
* ``fht_space_interpolation`` follows :235-292: tiled when either side exceeds the tile size (0 means 256).
"""
from __future__ import annotations

import math

import numpy as np


def _mirror_pad(img: np.ndarray) -> np.ndarray:
    h, w = img.shape
    bw = int(2 ** math.ceil(math.log2(w))) if w > 1 else 1
    bh = int(2 ** math.ceil(math.log2(h))) if h > 1 else 1
    n = max(bh, bw)
    sw = math.ceil(n / w)
    sw += 1 if sw % 2 == 0 else 0
    sh = math.ceil(n / h)
    sh += 1 if sh % 2 == 0 else 0
    mw, mh = (sw - 1) // 2, (sh - 1) // 2
    padded = np.zeros((sh * h, sw * w), dtype=np.float32)
    for tw in range(sw):
        cols = np.arange(w) if (tw - mw) % 2 == 0 else np.arange(w - 1, -1, -1)
        for th in range(sh):
            rows = np.arange(h) if (th - mh) % 2 == 0 else np.arange(h - 1, -1, -1)
            padded[th * h:(th + 1) * h, tw * w:(tw + 1) * w] = img[np.ix_(rows, cols)]
    xr = (sw * w - n) // 2
    yr = (sh * h - n) // 2
    return padded[yr:yr + n, xr:xr + n]


def fht_2d(img: np.ndarray, mag: int, mirror: bool = True) -> np.ndarray:
    img = np.asarray(img, dtype=np.float32)
    lo, hi = float(img.min()), float(img.max())
    work = _mirror_pad(img) if mirror else img
    h, w = work.shape
    hI, wI = h * mag, w * mag
    F = np.fft.rfft2(work)
    h2 = (h + 1) // 2
    Fi = np.zeros((hI, wI // 2 + 1), dtype=np.complex128)
    Fi[0:h2, 0:F.shape[1]] = F[0:h2, :]
    Fi[hI - (h - h2):hI, 0:F.shape[1]] = F[h2:h, :]
    out = np.fft.irfft2(Fi, s=(hI, wI)).astype(np.float32)
    omin, omax = float(out.min()), float(out.max())
    if omax != omin:
        scale = (hi - lo) / (omax - omin)
        out -= omin
        out *= scale
        out += lo
    if mirror:
        xr = (wI - mag * img.shape[1]) // 2
        yr = (hI - mag * img.shape[0]) // 2
        return out[yr:yr + mag * img.shape[0], xr:xr + mag * img.shape[1]]
    return out


def tile_starts(dim: int, tile: int, overlap: int) -> list:
    step = max(1, tile - 2 * overlap)
    s = list(range(0, dim, step))
    if s and s[-1] + tile < dim:
        s.append(max(0, dim - tile))
    return s


def fht_2d_tiled(img: np.ndarray, mag: int, tile: int = 256, overlap: int = 16) -> np.ndarray:
    img = np.asarray(img, dtype=np.float32)
    h, w = img.shape
    if h <= tile and w <= tile:
        return fht_2d(img, mag, True)
    out = np.zeros((h * mag, w * mag), dtype=np.float32)
    weight = np.zeros((h * mag, w * mag), dtype=np.float32)
    for r0 in tile_starts(h, tile, overlap):
        r1 = min(r0 + tile, h)
        for c0 in tile_starts(w, tile, overlap):
            c1 = min(c0 + tile, w)
            t = fht_2d(np.ascontiguousarray(img[r0:r1, c0:c1]), mag, True)
            th, tw = t.shape
            win = np.outer(np.hanning(th + 2)[1:-1].astype(np.float32), np.hanning(tw + 2)[1:-1].astype(np.float32))
            out[r0 * mag:r1 * mag, c0 * mag:c1 * mag] += t * win
            weight[r0 * mag:r1 * mag, c0 * mag:c1 * mag] += win
    nz = weight > 0
    out[nz] /= weight[nz]
    return out


def fht_space_interpolation(image, magnification: int = 2, doMirrorPadding: bool = True, tile_size: int = 0,
                            overlap: int = 16) -> np.ndarray:
    a = np.asarray(image)
    if a.ndim not in (2, 3):
        raise ValueError("Input must be 2-D or 3-D array")
    h, w = a.shape[-2:]
    tile = int(tile_size) if int(tile_size) > 0 else 256
    frames = a if a.ndim == 3 else a[None]
    if h > tile or w > tile:
        res = np.stack([fht_2d_tiled(np.ascontiguousarray(f, dtype=np.float32), int(magnification), tile,
                                     int(overlap)) for f in frames])
    else:
        res = np.stack([fht_2d(np.ascontiguousarray(f, dtype=np.float32), int(magnification), doMirrorPadding)
                        for f in frames])
    return res if a.ndim == 3 else res[0]


def esrrf_fht(stack, magnification=5, radius=1.5, sensitivity=1, doIntensityWeighting=True, pad_edges=False):
    """eSRRF._run_threaded with use_fht=True (src/nanopyx/core/transform/_le_esrrf.pyx:183-188): Fourier xM of the
    image, Roberts cross of the image, Fourier x2M of both gradients, RGC on those arrays with pad_edges forwarded."""
    from . import ref32
    stack = np.ascontiguousarray(stack if np.ndim(stack) == 3 else np.asarray(stack)[None], dtype=np.float32)
    M = int(magnification)
    mag = fht_space_interpolation(stack, M)
    gc, gr = ref32.roberts(stack)
    gci = fht_space_interpolation(gc, 2 * M)
    gri = fht_space_interpolation(gr, 2 * M)
    return ref32.rgc(gci, gri, mag, M, radius, sensitivity, doIntensityWeighting, pad_edges)
