"""Catmull-Rom ShiftAndMagnify operator (exported as CRShiftAndMagnify) with the gfx950 run type: drop-in for
src/nanopyx/core/transform/_le_interpolation_catmull_rom.pyx (``run(image, shift_row, shift_col, magnification_row,
magnification_col, run_type=None)``, :27-46, returning the (N, int(H * mag_row), int(W * mag_col)) float32 stack,
:143-170); one launch of libnpx_hip ``npx_shift_magnify``."""
from __future__ import annotations

from ... import _hip
from ...__liquid_engine__ import LiquidEngine
from ._le_esrrf import _is_tensor
from .__interpolation_tools__ import check_image


class ShiftAndMagnify(LiquidEngine):
    """Shift and magnify using Catmull-Rom interpolation and the Liquid Engine."""

    def __init__(self, clear_benchmarks=False, testing=False, verbose=True):
        self._designation = "ShiftMagnify_catmull_rom"
        super().__init__(clear_benchmarks=clear_benchmarks, testing=testing, verbose=verbose)

    def run(self, image, shift_row, shift_col, magnification_row: float, magnification_col: float, run_type=None):
        image = check_image(image, keep_uint16=True)
        return self._run(image, shift_row, shift_col, magnification_row, magnification_col, run_type=run_type)

    def benchmark(self, image, shift_row, shift_col, magnification_row: float, magnification_col: float):
        image = check_image(image, keep_uint16=True)
        return super().benchmark(image, shift_row, shift_col, magnification_row, magnification_col)

    def _run_hip(self, image, shift_row, shift_col, magnification_row, magnification_col):
        """
        @gpu
        """
        out = _hip.shift_magnify(image, float(shift_row), float(shift_col), float(magnification_row),
                                 float(magnification_col))
        return out if _is_tensor(image) else out.cpu().numpy()
