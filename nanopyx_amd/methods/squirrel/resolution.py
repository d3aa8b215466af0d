"""``calculate_frc`` and ``calculate_decorr_analysis`` (src/nanopyx/methods/squirrel/resolution.py:7-110): the FIRE
number of two frames and the decorrelation resolution of one frame, on the host."""
from __future__ import annotations

import numpy as np

from ...core.analysis.decorr import DecorrAnalysis
from ...core.analysis.frc import FIRECalculator


def calculate_frc(frame_1: np.ndarray, frame_2: np.ndarray, pixel_size: float = 1, units: str = "pixel",
                  plot_frc_curve: bool = False):
    if plot_frc_curve:
        raise NotImplementedError("plot_frc_curve is not part of nanopyx_amd")
    return FIRECalculator(pixel_size=pixel_size, units=units).calculate_fire_number(np.asarray(frame_1),
                                                                                     np.asarray(frame_2))


def calculate_decorr_analysis(frame: np.ndarray, rmin: float = 0, rmax: float = 1, n_r: int = 50, n_g: int = 10,
                              pixel_size: float = 1, units: str = "pixel", roi: tuple = (0, 0, 0, 0),
                              plot_decorr_analysis=False):
    if plot_decorr_analysis:
        raise NotImplementedError("plot_decorr_analysis is not part of nanopyx_amd")
    d = DecorrAnalysis(pixel_size=pixel_size, units=units, rmin=rmin, rmax=rmax, n_r=n_r, n_g=n_g, roi=roi)
    d.run_analysis(frame)
    return d.resolution
