"""Radiality operator (SRRF) with the gfx950 run type.

Drop-in for src/nanopyx/core/transform/_le_radiality.pyx:27-363: same constructor and ``run(image, image_interp,
magnification=5, ringRadius=0.5, border=0, radialityPositivityConstraint=True, doIntensityWeighting=True,
run_type=None)`` (:40-43), returning the (N, H*M, W*M) float32 radiality maps. The HIP_gfx950 run type is one launch
of libnpx_hip ``npx_radiality`` over the whole stack (gradients, the 12-sample Catmull-Rom ring and the intensity
weighting fused per frame); ``image_interp`` is the caller's xM image, read only when ``doIntensityWeighting``.
"""
from __future__ import annotations

import time

from ... import _hip
from ...__liquid_engine__ import LiquidEngine
from ._le_esrrf import _is_tensor
from .__interpolation_tools__ import check_image


class Radiality(LiquidEngine):
    """Radiality (SRRF) using the Liquid Engine."""

    def __init__(self, clear_benchmarks=False, testing=False, verbose=True):
        self._designation = "Radiality"
        super().__init__(clear_benchmarks=clear_benchmarks, testing=testing, verbose=verbose)

    def run(self, image, image_interp, magnification: int = 5, ringRadius: float = 0.5, border: int = 0,
            radialityPositivityConstraint: bool = True, doIntensityWeighting: bool = True, run_type=None):
        image = check_image(image, keep_uint16=True)
        image_interp = check_image(image_interp)
        return self._run(image, image_interp, magnification, ringRadius, border, radialityPositivityConstraint,
                         doIntensityWeighting, run_type=run_type)

    def benchmark(self, image, image_interp, magnification: int = 5, ringRadius: float = 0.5, border: int = 0,
                  radialityPositivityConstraint: bool = True, doIntensityWeighting: bool = True):
        image = check_image(image, keep_uint16=True)
        image_interp = check_image(image_interp)
        return super().benchmark(image, image_interp, magnification, ringRadius, border,
                                 radialityPositivityConstraint, doIntensityWeighting)

    def _run_hip(self, image, image_interp, magnification=5, ringRadius=0.5, border=0,
                 radialityPositivityConstraint=True, doIntensityWeighting=True):
        """
        @gpu
        """
        out = _hip.radiality(image, image_interp if doIntensityWeighting else None, int(magnification), ringRadius,
                             int(border), radialityPositivityConstraint, doIntensityWeighting)
        return out if _is_tensor(image) else out.cpu().numpy()

    def run_fused(self, image, magnification: int = 5, ringRadius: float = 0.5, border: int = 0,
                  radialityPositivityConstraint: bool = True, doIntensityWeighting: bool = True, stream=None):
        """CRShiftAndMagnify + Radiality in one launch (the intensity image is magnified inside the kernel);
        returns the device tensor."""
        image = check_image(image, keep_uint16=True)
        t0 = time.perf_counter()
        out = _hip.radiality(image, None, int(magnification), ringRadius, int(border), radialityPositivityConstraint,
                             doIntensityWeighting, stream=stream)
        self._last_runtype, self._last_time = "HIP_gfx950", time.perf_counter() - t0
        return out
