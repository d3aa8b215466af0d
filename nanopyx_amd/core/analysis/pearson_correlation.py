"""Pearson correlation of two images (src/nanopyx/core/analysis/pearson_correlation.pyx:40-77).

The reference accumulates in float32 over an OpenMP-parallel loop (order not fixed); this restatement accumulates
in float64 and returns the float32 result, 0 when either image is constant."""
from __future__ import annotations

import numpy as np


def pearson_correlation(im1, im2) -> float:
    a = np.asarray(im1, dtype=np.float32).astype(np.float64)
    b = np.asarray(im2, dtype=np.float32).astype(np.float64)
    da = a - a.mean()
    db = b - b.mean()
    s11 = float((da * da).sum())
    s22 = float((db * db).sum())
    if s11 == 0 or s22 == 0:
        return 0.0
    return float(np.float32(float((da * db).sum()) / (s11 * s22) ** 0.5))
