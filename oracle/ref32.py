"""ORACLE (test infrastructure only) — ctypes front end of the FP32 restatement of the reference CPU path.

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may import this module.

``libnpx_oracle.so`` (oracle/esrrf_ref32.c, oracle/post_ref32.c) restates the reference's threaded Cython
eSRRF chain stage for stage (see the C file for the file:line each function follows). With
``use_reference_kernels(True)`` its loop drivers call the reference's own per-pixel C kernels, compiled
from the reference sources into ``oracle/_ref/libnpx_ref.so`` by ``oracle/Makefile``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libnpx_oracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libnpx_ref.so")
_REF_SRRF_PATH = os.path.join(_HERE, "_ref", "libnpx_ref_srrf.so")
_lib = None
_ref = None
_ref_srrf = None

_fp = ctypes.POINTER(ctypes.c_float)


def build(force: bool = False) -> None:
    """Compile the oracle (and oracle/_ref when the reference sources are present) with make."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "-s"], check=True,
                       stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.ora_shift_magnify.argtypes = [_fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                        ctypes.c_float, ctypes.c_float, ctypes.c_float, _fp]
        L.ora_roberts.argtypes = [_fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp, _fp]
        L.ora_rgc.argtypes = [_fp, _fp, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, _fp]
        L.ora_esrrf.argtypes = [_fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                ctypes.c_double, ctypes.c_int, ctypes.c_int, _fp]
        L.ora_esrrf.restype = ctypes.c_int
        L.ora_temporal.argtypes = [_fp, ctypes.c_int, ctypes.c_long, ctypes.c_int, _fp]
        L.ora_macro_pixel.argtypes = [_fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ora_use_kernels.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.ora_px_interpolate.argtypes = [_fp, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int]
        L.ora_px_interpolate.restype = ctypes.c_float
        L.ora_px_rgc.argtypes = [ctypes.c_int, ctypes.c_int, _fp, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.ora_px_rgc.restype = ctypes.c_float
        L.ora_px_roberts.argtypes = [_fp, _fp, _fp, ctypes.c_int, ctypes.c_int]
        L.ora_num_threads.restype = ctypes.c_int
        L.ora_use_3d_kernels.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.ora_esrrf3d.argtypes = [_fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_int, ctypes.c_int, _fp]
        L.ora_esrrf3d.restype = ctypes.c_int
        L.ora_use_radiality_kernels.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.ora_px_radiality.argtypes = [ctypes.c_int, ctypes.c_int, _fp, _fp, _fp, _fp, ctypes.c_int, ctypes.c_float,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ora_px_radiality.restype = ctypes.c_float
        L.ora_px_gradient_radiality.argtypes = [_fp, _fp, _fp, ctypes.c_int, ctypes.c_int]
        L.ora_radiality_ring.argtypes = [ctypes.c_double, ctypes.c_int, _fp, _fp, _fp]
        L.ora_radiality.argtypes = [_fp, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp]
        _lib = L
    return _lib


def reference_available() -> bool:
    return os.path.exists(_REF_PATH)


def use_reference_kernels(on: bool) -> None:
    """Route the per-pixel kernels to the reference's compiled C (True) or to the restatement (False)."""
    global _ref
    L = lib()
    if not on:
        L.ora_use_kernels(None, None, None)
        return
    if _ref is None:
        _ref = ctypes.CDLL(_REF_PATH)
    addr = lambda name: ctypes.cast(getattr(_ref, name), ctypes.c_void_p).value
    L.ora_use_kernels(addr("_c_calculate_rgc"), addr("_c_interpolate"), addr("_c_gradient_roberts_cross"))


def use_reference_radiality(on: bool) -> None:
    """Route the radiality kernels to the reference's compiled _c_calculate_radiality_per_subpixel /
    _c_gradient_radiality (True) or to the restatement (False)."""
    global _ref_srrf
    L = lib()
    if not on:
        L.ora_use_radiality_kernels(None, None)
        return
    if _ref_srrf is None:
        _ref_srrf = ctypes.CDLL(_REF_SRRF_PATH)
    addr = lambda name: ctypes.cast(getattr(_ref_srrf, name), ctypes.c_void_p).value
    L.ora_use_radiality_kernels(addr("_c_calculate_radiality_per_subpixel"), addr("_c_gradient_radiality"))


def use_reference_3d(on: bool) -> None:
    """Route the eSRRF 3D kernels to the reference's compiled _c_calculate_rgc3D / _c_gradient_3d (True) or to the
    restatement (False); the xy Catmull-Rom follows use_reference_kernels."""
    global _ref
    L = lib()
    if not on:
        L.ora_use_3d_kernels(None, None)
        return
    if _ref is None:
        _ref = ctypes.CDLL(_REF_PATH)
    addr = lambda name: ctypes.cast(getattr(_ref, name), ctypes.c_void_p).value
    L.ora_use_3d_kernels(addr("_c_calculate_rgc3D"), addr("_c_gradient_3d"))


def reference_srrf_available() -> bool:
    return os.path.exists(_REF_SRRF_PATH)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a):
    return a.ctypes.data_as(_fp)


def shift_magnify(img, mag_row, mag_col, shift_row=0.0, shift_col=0.0):
    img = _f32(img if np.ndim(img) == 3 else np.asarray(img)[None])
    N, H, W = img.shape
    out = np.zeros((N, int(H * np.float32(mag_row)), int(W * np.float32(mag_col))), np.float32)
    lib().ora_shift_magnify(_p(img), N, H, W, shift_row, shift_col, mag_row, mag_col, _p(out))
    return out


def roberts(img):
    img = _f32(img if np.ndim(img) == 3 else np.asarray(img)[None])
    gc = np.zeros_like(img)
    gr = np.zeros_like(img)
    lib().ora_roberts(_p(img), *img.shape, _p(gc), _p(gr))
    return gc, gr


def rgc(gci, gri, img_i, magnification=5, radius=1.5, sensitivity=1, doIntensityWeighting=True, pad_edges=False):
    gci, gri, img_i = _f32(gci), _f32(gri), _f32(img_i)
    N, HM, WM = img_i.shape
    out = np.zeros_like(img_i)
    lib().ora_rgc(_p(gci), _p(gri), _p(img_i), N, HM, WM, int(magnification), float(radius), float(sensitivity),
                  int(bool(doIntensityWeighting)), int(bool(pad_edges)), _p(out))
    return out


def esrrf(stack, magnification=5, radius=1.5, sensitivity=1, doIntensityWeighting=True, chunk=0):
    """eSRRF_ST.run on the CPU ("Threaded" chain), (N, HM, WM) float32."""
    stack = _f32(stack if np.ndim(stack) == 3 else np.asarray(stack)[None])
    N, H, W = stack.shape
    M = int(magnification)
    out = np.zeros((N, H * M, W * M), np.float32)
    rc = lib().ora_esrrf(_p(stack), N, H, W, M, float(radius), float(sensitivity), int(bool(doIntensityWeighting)),
                         int(chunk), _p(out))
    if rc != 0:
        raise MemoryError("oracle eSRRF could not allocate its intermediates")
    return out


_MODES = {"AVG": 0, "VAR": 1, "TAC2": 2}


def temporal(stack, mode):
    stack = _f32(stack)
    N = stack.shape[0]
    P = int(np.prod(stack.shape[1:]))
    out = np.zeros(stack.shape[1:], np.float32)
    lib().ora_temporal(_p(stack), N, P, _MODES[mode], _p(out))
    return out


def macro_pixel(img, magnification):
    x = np.array(img, dtype=np.float32, copy=True)
    if x.ndim == 2:
        x = x[None]
    lib().ora_macro_pixel(_p(x), x.shape[0], x.shape[1], x.shape[2], int(magnification))
    return x


def esrrf_workflow(image, magnification=5, radius=1.5, sensitivity=1, frames_per_timepoint=0,
                   temporal_correlation="AVG", doIntensityWeighting=True, macro_pixel_correction=True, chunk=0):
    """eSRRF() of src/nanopyx/methods/esrrf/eSRRF_workflow.py:12-121 over the FP32 chain."""
    image = np.asarray(image)
    N = image.shape[0]
    fpt = N if frames_per_timepoint == 0 or frames_per_timepoint > N else frames_per_timepoint
    T = N // fpt + (1 if N % fpt else 0)
    out = np.zeros((T, image.shape[1] * magnification, image.shape[2] * magnification), np.float32)
    for t in range(T):
        rgc_maps = esrrf(image[fpt * t:fpt * (t + 1)], magnification, radius, sensitivity, doIntensityWeighting,
                         chunk)
        m = temporal(rgc_maps, temporal_correlation)[None]
        out[t] = macro_pixel(m, magnification)[0] if macro_pixel_correction else m[0]
    return np.squeeze(out.astype(np.float32))


def radiality_ring(ringRadius, magnification):
    """(xRing[12], yRing[12], ringRadius * M) of Radiality._run_threaded."""
    x = np.zeros(12, np.float32)
    y = np.zeros(12, np.float32)
    r = np.zeros(1, np.float32)
    lib().ora_radiality_ring(float(ringRadius), int(magnification), _p(x), _p(y), _p(r))
    return x, y, float(r[0])


def radiality(image, image_interp, magnification=5, ringRadius=0.5, border=0, radialityPositivityConstraint=True,
              doIntensityWeighting=True):
    """Radiality.run on the CPU ("Threaded"), (N, H*M, W*M) float32."""
    image = _f32(image if np.ndim(image) == 3 else np.asarray(image)[None])
    image_interp = _f32(image_interp if np.ndim(image_interp) == 3 else np.asarray(image_interp)[None])
    N, H, W = image.shape
    M = int(magnification)
    assert image_interp.shape == (N, H * M, W * M)
    out = np.zeros((N, H * M, W * M), np.float32)
    lib().ora_radiality(_p(image), _p(image_interp), N, H, W, M, float(ringRadius), int(border),
                        int(bool(radialityPositivityConstraint)), int(bool(doIntensityWeighting)), _p(out))
    return out


def srrf_workflow(image, magnification=5, ringRadius=0.5, border=0, radialityPositivityConstraint=True,
                  doIntensityWeighting=True, macro_pixel_correction=True):
    """SRRF() of src/nanopyx/methods/srrf/SRRF_workflow.py:9-76: CR xM magnification, radiality, macro-pixel
    correction of every frame."""
    image = _f32(image if np.ndim(image) == 3 else np.asarray(image)[None])
    M = int(magnification)
    interp = shift_magnify(image, M, M)
    rad = radiality(image, interp, M, ringRadius, border, radialityPositivityConstraint, doIntensityWeighting)
    return macro_pixel(rad, M) if macro_pixel_correction else rad


_MODES3D = {"average": 0, "std": 1}


def esrrf3d(image, magnification_xy=2, magnification_z=2, radius=1.5, PSF_ratio=2.8, voxel_ratio=4.0,
            sensitivity=1, mode="average", doIntensityWeighting=True):
    """eSRRF3D.run on the CPU ("Threaded"): (slices*Mz, rows*Mxy, cols*Mxy) float32 for an (F, S, R, C) or (S, R, C)
    stack; radius_z = radius * PSF_ratio / voxel_ratio as in src/nanopyx/core/transform/_le_esrrf3d.pyx:40."""
    image = _f32(image if np.ndim(image) == 4 else np.asarray(image)[None])
    F, S, R, C = image.shape
    radius_z = radius * PSF_ratio / voxel_ratio
    out = np.zeros((S * int(magnification_z), R * int(magnification_xy), C * int(magnification_xy)), np.float32)
    rc = lib().ora_esrrf3d(_p(image), F, S, R, C, int(magnification_xy), int(magnification_z), float(radius),
                           float(radius_z), float(voxel_ratio), float(sensitivity), _MODES3D[mode],
                           int(bool(doIntensityWeighting)), _p(out))
    if rc != 0:
        raise MemoryError("oracle eSRRF 3D could not allocate its intermediates")
    return out


def num_threads() -> int:
    return lib().ora_num_threads()
