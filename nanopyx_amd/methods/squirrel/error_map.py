"""``calculate_error_map`` (src/nanopyx/methods/squirrel/error_map.py:5-36): error map, RSE and RSP of a
super-resolved image against a diffraction-limited reference, on the host."""
from __future__ import annotations

import numpy as np

from ...core.transform.error_map import ErrorMap


def calculate_error_map(img_ref: np.ndarray, img_sr: np.ndarray):
    emc = ErrorMap()
    emc.optimise(np.asarray(img_ref), np.asarray(img_sr))
    return np.asarray(emc.imRSE, dtype=np.float32), emc.getRSE(), emc.getRSP()
