"""Writes tests/golden/fht_golden.npz: input/output vectors of NanoPyx's Fourier interpolation (the use_fht branch).

The outputs come from the reference's own pure-NumPy module src/nanopyx/core/transform/interpolation.py, loaded by
file path (it imports only math, os, concurrent.futures and NumPy; the nanopyx package itself is not imported), on
non-constant frames: whole frames with and without mirror padding, and tiled frames whose tiles overlap and end
ragged (Hann-window blending). The SHA-256 of the module is stored with the vectors.

Run from the repository root with the reference present:  python tests/golden/make_fht_golden.py
"""
from __future__ import annotations

import hashlib
import importlib.util
import os

import numpy as np

SRC = "/root/reference/src/nanopyx/core/transform/interpolation.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fht_golden.npz")

# name, (N, H, W), magnification, mirror padding, tile size (0: whole frame), overlap
CASES = [
    ("whole_mirror_m4", (2, 20, 28), 4, True, 0, 16),
    ("whole_nopad_m3", (1, 16, 24), 3, False, 0, 16),
    ("tiled_m3_t32", (2, 72, 56), 3, True, 32, 4),
    ("tiled_m2_t16", (1, 40, 47), 2, True, 16, 3),
]


def _frames(shape, seed):
    """Blobs on a smooth background plus noise: non-constant, band-rich frames."""
    rng = np.random.default_rng(seed)
    N, H, W = shape
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    out = np.empty(shape, np.float32)
    for f in range(N):
        img = 50.0 + 0.8 * xx + 0.5 * yy
        for _ in range(max(3, H * W // 60)):
            cy, cx, s, a = rng.uniform(0, H), rng.uniform(0, W), rng.uniform(0.8, 2.5), rng.uniform(50, 400)
            img += a * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * s * s))
        img += rng.normal(0, 5.0, size=img.shape)
        out[f] = img.astype(np.float32)
    return out


def main():
    spec = importlib.util.spec_from_file_location("reference_interpolation", SRC)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    data = {"source_sha256": np.frombuffer(hashlib.sha256(open(SRC, "rb").read()).hexdigest().encode(), np.uint8)}
    for k, (name, shape, M, mirror, tile, overlap) in enumerate(CASES):
        x = _frames(shape, 100 + k)
        y = mod.fht_space_interpolation(x, M, mirror, tile, overlap)
        data[f"{name}/input"] = x
        data[f"{name}/output"] = np.asarray(y, np.float32)
        data[f"{name}/params"] = np.array([M, int(mirror), tile, overlap], np.int32)
        print(name, x.shape, "->", y.shape, float(np.abs(y).max()))
    np.savez_compressed(OUT, **data)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
