"""Writes tests/golden/esrrf_golden.npz: input/output vectors of NanoPyx's eSRRF path.

The per-pixel arithmetic comes from the reference's own C kernels (_c_gradient_roberts_cross, _c_interpolate,
_c_calculate_rgc), compiled from the reference sources into oracle/_ref/libnpx_ref.so by oracle/Makefile and
driven by the restated Cython loops of oracle/esrrf_ref32.c; the temporal step is NumPy itself
(np.mean / np.var / calculate_tac2's expression on float32, src/nanopyx/core/transform/sr_temporal_correlations.py).
The SHA-256 of the reference C sources used is stored with the vectors.

Run from the repository root with the reference present:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import ref32  # noqa: E402
from nanopyx_amd import synthetic  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "esrrf_golden.npz")

CASES = [
    # name, (N, H, W), M, R, S, intensity weighting, input kind
    ("m5_r15_s1", (3, 12, 14), 5, 1.5, 1.0, True, "scene"),
    ("m4_r20_s2", (3, 16, 16), 4, 2.0, 2.0, True, "scene_u16"),
    ("m2_r10_s3_noiw", (2, 10, 12), 2, 1.0, 3.0, False, "random"),
    ("m3_r25_s17", (2, 9, 9), 3, 2.5, 1.7, True, "random"),
    ("m8_r12_s2", (2, 6, 7), 8, 1.2, 2.0, True, "scene"),
    ("m1_r15_s1", (2, 8, 8), 1, 1.5, 1.0, True, "random"),
]


def _input(kind, shape, seed):
    N, H, W = shape
    if kind == "random":
        rng = np.random.default_rng(seed)
        return (rng.random(shape) * 200 + 20).astype(np.float32)
    sc = synthetic.Scene(N, H, W, max(2, H * W // 20), 1.5, 600, 0.4, 100, 0.0, 2.0,
                         100.0 if kind == "scene_u16" else 0.0, "uint16" if kind == "scene_u16" else "float32", seed)
    return synthetic.generate(sc)


def main():
    if not ref32.reference_available():
        raise SystemExit("oracle/_ref/libnpx_ref.so missing: run `make -C oracle` with the reference sources present")
    ref32.use_reference_kernels(True)
    data = {}
    for k, (name, shape, M, R, S, iw, kind) in enumerate(CASES):
        x = _input(kind, shape, 100 + k)
        xf = x.astype(np.float32)
        gc, gr = ref32.roberts(xf)
        mag = ref32.shift_magnify(xf, float(M), float(M))
        gci = ref32.shift_magnify(gc, float(2 * M), float(2 * M))
        gri = ref32.shift_magnify(gr, float(2 * M), float(2 * M))
        rgc = ref32.rgc(gci, gri, mag, M, R, S, iw)
        maps = ref32.esrrf(xf, M, R, S, iw)
        assert np.array_equal(maps, rgc)
        data[f"{name}/input"] = x
        data[f"{name}/params"] = np.array([M, R, S, float(iw)], dtype=np.float64)
        data[f"{name}/roberts_gc"] = gc
        data[f"{name}/roberts_gr"] = gr
        data[f"{name}/magnified"] = mag
        data[f"{name}/rgc_maps"] = maps
        # gradient images at x2M are large; keep them for the smallest case only (RGC-class input)
        if name == "m2_r10_s3_noiw":
            data[f"{name}/grad_col_interp"] = gci
            data[f"{name}/grad_row_interp"] = gri
        maps32 = np.array(maps, dtype="float32")
        data[f"{name}/AVG"] = np.mean(maps32, axis=0)
        data[f"{name}/VAR"] = np.var(maps32, axis=0)
        centered = maps32 - np.mean(maps32, axis=0)
        data[f"{name}/TAC2"] = np.mean(centered[:-1] * centered[1:], axis=0)
    # RGC class with pad_edges=True on the stored x2M gradients
    name = "m2_r10_s3_noiw"
    data[f"{name}/rgc_pad_edges"] = ref32.rgc(data[f"{name}/grad_col_interp"], data[f"{name}/grad_row_interp"],
                                              data[f"{name}/magnified"], 2, 1.0, 3.0, True, pad_edges=True)
    ref32.use_reference_kernels(False)
    sha = open(os.path.join(ROOT, "oracle", "_ref", "SOURCES.sha256")).read()
    data["reference_sources_sha256"] = np.array(sha)
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes, {len(CASES)} cases)")


if __name__ == "__main__":
    main()
