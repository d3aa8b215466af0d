"""GradientRobertsCross operator with the gfx950 run type: drop-in for
src/nanopyx/core/transform/_le_roberts_cross_gradients.pyx (``run(image, run_type=None)`` returning the column and
row gradients of an (N, H, W) stack, :67-83), one launch of libnpx_hip ``npx_roberts``."""
from __future__ import annotations

from ... import _hip
from ...__liquid_engine__ import LiquidEngine
from ._le_esrrf import _is_tensor
from .__interpolation_tools__ import check_image


class GradientRobertsCross(LiquidEngine):
    """Roberts cross gradients using the Liquid Engine."""

    def __init__(self, clear_benchmarks=False, testing=False, verbose=True):
        self._designation = "GradientRobertsCross"
        super().__init__(clear_benchmarks=clear_benchmarks, testing=testing, verbose=verbose)

    def run(self, image, run_type=None):
        image = check_image(image, keep_uint16=True)
        return self._run(image, run_type=run_type)

    def benchmark(self, image):
        image = check_image(image, keep_uint16=True)
        return super().benchmark(image)

    def _run_hip(self, image):
        """
        @gpu
        """
        gc, gr = _hip.roberts(image)
        return (gc, gr) if _is_tensor(image) else (gc.cpu().numpy(), gr.cpu().numpy())
