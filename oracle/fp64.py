"""ORACLE (test infrastructure only) — FP64 NumPy statement of the eSRRF formulas.

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may import this module;
the product package never does.

This is the "FP64 CPU oracle": the reference's eSRRF chain evaluated in exact float64 arithmetic, with
exact rational sampling coordinates (``i / M`` and ``(xM + 0.5) / M`` in float64 instead of float32).
The radius-derived constants (sigma, tSS, tSO, the window half-width and the margin) are *defined* by
the reference in float32, so they are taken from ``ref_params`` exactly as the reference derives them.

Reference semantics followed (paths relative to the reference repo root):
  * Roberts cross ........ src/include/_c_gradients.c:53-72
  * Catmull-Rom .......... src/include/_c_interpolation_catmull_rom.c:6-52,
                           src/nanopyx/core/transform/_le_interpolation_catmull_rom.pyx:143-170
  * RGC ring sum ......... src/include/_c_sr_radial_gradient_convergence.c:4-90,
                           src/nanopyx/core/transform/_le_radial_gradient_convergence.pyx:79-115
  * eSRRF composition .... src/nanopyx/core/transform/_le_esrrf.pyx:172-195
  * temporal ............. src/nanopyx/core/transform/sr_temporal_correlations.py:46-75,205-221
  * macro-pixel .......... src/nanopyx/core/transform/mpcorrector.pyx:7-73
  * workflow ............. src/nanopyx/methods/esrrf/eSRRF_workflow.py:58-121

How it is pinned (tests/test_oracle.py): against the golden vectors made from the reference's own C
kernels, against the bit-exact FP32 restatement (oracle/esrrf_ref32.c) on random scenes, and against
closed forms and invariants (CR at half-integer coordinates, linear ramps, constant images, transpose
equivariance, scale equivariance, brute-force per-pixel sums on tiny inputs).
"""
from __future__ import annotations

import math

import numpy as np


# ----------------------------------------------------------------------------- parameters

def ref_params(radius: float, magnification: int, pad_edges: bool = False):
    """Radius-derived constants exactly as RadialGradientConvergence derives them
    (_le_radial_gradient_convergence.pyx:85-98) and the window of _c_calculate_rgc (:27-28)."""
    sigma = np.float32(radius / 2.355)
    fwhm = np.float32(radius)
    tSS = np.float32(np.float32(2) * sigma * sigma)
    tSO = np.float32(np.float32(2) * sigma + np.float32(1))
    margin = 0 if pad_edges else int(float(fwhm) * magnification)
    start = -int(np.float32(2.0) * fwhm)
    end = int(np.float32(2.0) * fwhm + np.float32(1))
    return dict(sigma=float(sigma), fwhm=float(fwhm), tSS=float(tSS), tSO=float(tSO), margin=margin,
                start=start, end=end)


# ----------------------------------------------------------------------------- Roberts cross

def roberts(img: np.ndarray):
    """(Gc, Gr) of one (H, W) frame, _c_gradients.c:53-72 (previous row/column clamped at 0)."""
    a = np.asarray(img, dtype=np.float64)
    up = np.concatenate([a[:1], a[:-1]], axis=0)            # I[r0, c1]
    left = np.concatenate([a[:, :1], a[:, :-1]], axis=1)    # I[r1, c0]
    ul = np.concatenate([up[:, :1], up[:, :-1]], axis=1)    # I[r0, c0]
    gc = up - left + a - ul
    gr = -up + left + a - ul
    return gc, gr


# ----------------------------------------------------------------------------- Catmull-Rom

def cubic(v):
    """Keys cubic with a = 0.5, _c_interpolation_catmull_rom.c:6-18."""
    v = np.abs(np.asarray(v, dtype=np.float64))
    a = 0.5
    inner = v * v * (v * (-a + 2) + (a - 3)) + 1
    outer = -a * v ** 3 + 5 * a * v * v - 8 * a * v + 4 * a
    return np.where(v < 1, inner, np.where(v < 2, outer, 0.0))


def cr_weight_matrix(coords, n: int) -> np.ndarray:
    """Row-weight matrix W (len(coords), n) such that W @ img samples the image at `coords` along one
    axis with the 4-tap rule of _c_interpolate: taps floor(x-0.5)-1 .. +2, taps outside [0, n) skipped,
    and an all-zero row when x < 0 or x >= n."""
    x = np.asarray(coords, dtype=np.float64)
    base = np.floor(x - 0.5).astype(np.int64) - 1
    Wm = np.zeros((x.size, n), dtype=np.float64)
    inside = (x >= 0) & (x < n)
    for u in range(4):
        idx = base + u
        w = cubic(x - (idx + 0.5))
        ok = inside & (idx >= 0) & (idx < n)
        rows = np.nonzero(ok)[0]
        np.add.at(Wm, (rows, idx[ok]), w[ok])
    return Wm


def cr_sample(img: np.ndarray, rcoords, ccoords) -> np.ndarray:
    """Separable Catmull-Rom of one frame on the grid rcoords x ccoords (float64)."""
    a = np.asarray(img, dtype=np.float64)
    Wr = cr_weight_matrix(rcoords, a.shape[0])
    Wc = cr_weight_matrix(ccoords, a.shape[1])
    return Wr @ a @ Wc.T


def magnify(img: np.ndarray, M: int) -> np.ndarray:
    """ShiftAndMagnify with zero shift at exact coordinates i / M."""
    H, W = img.shape
    return cr_sample(img, np.arange(H * M) / M, np.arange(W * M) / M)


# ----------------------------------------------------------------------------- RGC

def _lattice_gradients(gc, gr, H, W, M, n_a, n_b):
    """Interpolated gradients at lattice points a, b >= 0 (positions a/2, b/2), i.e. the values RGC reads
    from the 2M-upsampled gradient images at pos = a*M, clamped to the last row/column 2*H*M-1."""
    pa = np.minimum(np.arange(n_a) * M, 2 * H * M - 1)
    pb = np.minimum(np.arange(n_b) * M, 2 * W * M - 1)
    rc = pa / (2.0 * M)
    cc = pb / (2.0 * M)
    return cr_sample(gc, rc, cc), cr_sample(gr, rc, cc)


def rgc_frame_from_lattice(GX, GY, HM, WM, M, radius, sensitivity, pad_edges=False):
    """RGC ring sum of _c_calculate_rgc for every pixel of one frame (no intensity weighting), with the
    interpolated gradients given on the half-pixel lattice: GX[a, b] is the gradient at (a/2, b/2)."""
    p = ref_params(radius, M, pad_edges)
    tSS, tSO, margin = p["tSS"], p["tSO"], p["margin"]
    out = np.zeros((HM, WM), dtype=np.float64)
    r = np.arange(margin, HM - margin)
    c = np.arange(margin, WM - margin)
    if r.size == 0 or c.size == 0:
        return out
    yc = (r + 0.5) / M
    xc = (c + 0.5) / M
    a0 = np.floor(2 * yc).astype(np.int64)
    b0 = np.floor(2 * xc).astype(np.int64)
    num = np.zeros((r.size, c.size))
    den = np.zeros((r.size, c.size))
    for j in range(p["start"], p["end"]):
        a = a0 + j
        vy = a / 2.0
        oky = (vy > 0) & (vy <= HM - 1)
        dy = vy - yc
        for i in range(p["start"], p["end"]):
            b = b0 + i
            vx = b / 2.0
            okx = (vx > 0) & (vx <= WM - 1)
            dx = vx - xc
            d = np.sqrt(dy[:, None] ** 2 + dx[None, :] ** 2)
            ok = oky[:, None] & okx[None, :] & (d != 0) & (d <= tSO)
            if not ok.any():
                continue
            ai = np.clip(a, 0, GX.shape[0] - 1)
            bi = np.clip(b, 0, GX.shape[1] - 1)
            gx = GX[ai[:, None], bi[None, :]]
            gy = GY[ai[:, None], bi[None, :]]
            dsafe = np.where(d == 0, 1.0, d)
            dw = (dsafe * np.exp(-dsafe * dsafe / tSS)) ** 4
            dw = np.where(ok, dw, 0.0)
            den += dw
            gdotr = gx * dx[None, :] + gy * dy[:, None]
            nrm = np.sqrt(gx * gx + gy * gy)
            with np.errstate(invalid="ignore", divide="ignore"):
                dk = np.abs(gy * dx[None, :] - gx * dy[:, None]) / nrm
            dk = np.where(np.isnan(dk), dsafe, dk)
            dk = 1.0 - dk / dsafe
            num += np.where(ok & (gdotr < 0), dk * dw, 0.0)
    with np.errstate(invalid="ignore", divide="ignore"):
        rgc = np.where(den > 0, num / np.where(den > 0, den, 1.0), 0.0)
    S = float(np.float32(sensitivity))
    rgc = np.where(rgc < 0, 0.0, rgc)
    if S > 1:
        rgc = rgc ** S
    out[margin:HM - margin, margin:WM - margin] = rgc
    return out


def esrrf_frames(stack: np.ndarray, magnification=5, radius=1.5, sensitivity=1, doIntensityWeighting=True):
    """eSRRF_ST.run (non-FHT): per-frame RGC (x intensity) maps, (N, H*M, W*M) float64."""
    stack = np.asarray(stack, dtype=np.float64)
    if stack.ndim == 2:
        stack = stack[None]
    N, H, W = stack.shape
    M = int(magnification)
    HM, WM = H * M, W * M
    p = ref_params(radius, M)
    K = -p["start"]
    # every lattice index a that any pixel can reach: floor(2*yc) + K <= 2*H + K
    n_a = 2 * H + K + 2
    n_b = 2 * W + K + 2
    out = np.zeros((N, HM, WM))
    for f in range(N):
        img = stack[f]
        gc, gr = roberts(img)
        GX, GY = _lattice_gradients(gc, gr, H, W, M, n_a, n_b)
        rgc = rgc_frame_from_lattice(GX, GY, HM, WM, M, radius, sensitivity)
        if doIntensityWeighting:
            rgc = rgc * magnify(img, M)
        out[f] = rgc
    return out


def rgc_from_arrays(gx_i, gy_i, img_i, magnification=5, radius=1.5, sensitivity=1, doIntensityWeighting=True,
                    pad_edges=False):
    """RadialGradientConvergence.run on caller arrays (N, 2HM, 2WM) x2 and (N, HM, WM), indexed at
    pos = vy * 2M (clamped) exactly as _c_calculate_rgc reads them."""
    gx_i = np.asarray(gx_i, dtype=np.float64)
    gy_i = np.asarray(gy_i, dtype=np.float64)
    img_i = np.asarray(img_i, dtype=np.float64)
    N, HM, WM = img_i.shape
    M = int(magnification)
    p = ref_params(radius, M, pad_edges)
    K = -p["start"]
    out = np.zeros((N, HM, WM))
    # lattice index a reads row min(a*M, 2*HM - 1) of the caller array (row width 2*WM)
    n_a = int(2 * HM // M + K + 4)
    n_b = int(2 * WM // M + K + 4)
    pa = np.minimum(np.arange(n_a) * M, 2 * HM - 1)
    pb = np.minimum(np.arange(n_b) * M, 2 * WM - 1)
    for f in range(N):
        flat_x = gx_i[f].reshape(-1)
        flat_y = gy_i[f].reshape(-1)
        idx = pa[:, None] * (2 * WM) + pb[None, :]
        GX = flat_x[np.minimum(idx, flat_x.size - 1)]
        GY = flat_y[np.minimum(idx, flat_y.size - 1)]
        rgc = rgc_frame_from_lattice(GX, GY, HM, WM, M, radius, sensitivity, pad_edges)
        if doIntensityWeighting:
            rgc = rgc * img_i[f]
        out[f] = rgc
    return out


# ----------------------------------------------------------------------------- temporal / post

def temporal(stack: np.ndarray, mode: str) -> np.ndarray:
    """AVG / VAR (ddof 0) / TAC2 over axis 0 in float64 (sr_temporal_correlations.py:46-75,205-221)."""
    x = np.asarray(stack, dtype=np.float64)
    if mode == "AVG":
        return x.mean(axis=0)
    if mode == "VAR":
        return x.var(axis=0)
    if mode == "TAC2":
        c = x - x.mean(axis=0)
        return (c[:-1] * c[1:]).mean(axis=0)
    raise ValueError("Type of correlation must be AVG, VAR or TAC2")


def macro_pixel(img: np.ndarray, M: int) -> np.ndarray:
    """macro_pixel_corrector (mpcorrector.pyx:7-73) in float64 on a (S, HM, WM) stack."""
    x = np.array(img, dtype=np.float64, copy=True)
    S, HM, WM = x.shape
    rows, cols = HM // M, WM // M
    for s in range(S):
        blk = x[s, :rows * M, :cols * M].reshape(rows, M, cols, M)
        cmap = blk.sum(axis=(0, 2)) / (rows * cols)
        tiled = np.tile(cmap, (math.ceil(HM / M), math.ceil(WM / M)))[:HM, :WM]
        nz = tiled != 0
        x[s][nz] = x[s][nz] / tiled[nz]
    return x


def esrrf_workflow(image, magnification=5, radius=1.5, sensitivity=1, frames_per_timepoint=0,
                   temporal_correlation="AVG", doIntensityWeighting=True, macro_pixel_correction=True):
    """eSRRF() of eSRRF_workflow.py:12-121 in float64: timepoints, temporal step, macro-pixel, squeeze."""
    image = np.asarray(image, dtype=np.float64)
    N = image.shape[0]
    fpt = N if frames_per_timepoint == 0 or frames_per_timepoint > N else frames_per_timepoint
    T = N // fpt + (1 if N % fpt else 0)
    out = np.zeros((T, image.shape[1] * magnification, image.shape[2] * magnification))
    for t in range(T):
        rgc = esrrf_frames(image[fpt * t:fpt * (t + 1)], magnification, radius, sensitivity, doIntensityWeighting)
        m = temporal(rgc, temporal_correlation)[None]
        out[t] = macro_pixel(m, magnification)[0] if macro_pixel_correction else m[0]
    return np.squeeze(out)
