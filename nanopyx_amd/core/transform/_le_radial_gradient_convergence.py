"""RadialGradientConvergence operator with the gfx950 run type.

Drop-in for src/nanopyx/core/transform/_le_radial_gradient_convergence.pyx:16-332: same constructor and
``run(gradient_col_interp, gradient_row_interp, image_interp, magnification=5, radius=1.5, sensitivity=1,
doIntensityWeighting=True, pad_edges=False, run_type=None)`` (:28-32). The inputs are arbitrary caller arrays,
(N, 2*rowsM, 2*colsM) gradients and an (N, rowsM, colsM) intensity, read exactly where _c_calculate_rgc reads
them (libnpx_hip ``npx_rgc``); ``pad_edges`` sets the margin to 0 (:92-95).
"""
from __future__ import annotations

from ... import _hip
from ...__liquid_engine__ import LiquidEngine
from ._le_esrrf import _is_tensor
from .__interpolation_tools__ import check_image


class RadialGradientConvergence(LiquidEngine):
    """Radial gradient convergence using the Liquid Engine."""

    def __init__(self, clear_benchmarks=False, testing=False, verbose=True):
        self._designation = "RadialGradientConvergence"
        super().__init__(clear_benchmarks=clear_benchmarks, testing=testing, verbose=verbose)

    def run(self, gradient_col_interp, gradient_row_interp, image_interp, magnification: int = 5,
            radius: float = 1.5, sensitivity: float = 1, doIntensityWeighting: bool = True, pad_edges: bool = False,
            run_type=None):
        gradient_col_interp = check_image(gradient_col_interp)
        gradient_row_interp = check_image(gradient_row_interp)
        image_interp = check_image(image_interp)
        return self._run(gradient_col_interp, gradient_row_interp, image_interp, magnification, radius, sensitivity,
                         doIntensityWeighting, pad_edges=pad_edges, run_type=run_type)

    def benchmark(self, gradient_col_interp, gradient_row_interp, image_interp, magnification: int = 5,
                  radius: float = 1.5, sensitivity: float = 1, doIntensityWeighting: bool = True,
                  pad_edges: bool = False):
        gradient_col_interp = check_image(gradient_col_interp)
        gradient_row_interp = check_image(gradient_row_interp)
        image_interp = check_image(image_interp)
        return super().benchmark(gradient_col_interp, gradient_row_interp, image_interp, magnification, radius,
                                 sensitivity, doIntensityWeighting, pad_edges=pad_edges)

    def _run_hip(self, gradient_col_interp, gradient_row_interp, image_interp, magnification=5, radius=1.5,
                 sensitivity=1, doIntensityWeighting=True, pad_edges=False):
        """
        @gpu
        """
        out = _hip.rgc(gradient_col_interp, gradient_row_interp, image_interp, int(magnification), radius,
                       sensitivity, doIntensityWeighting, pad_edges)
        return out if _is_tensor(image_interp) else out.cpu().numpy()
