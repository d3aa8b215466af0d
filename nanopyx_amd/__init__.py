"""nanopyx_amd — NanoPyx's eSRRF hot path on AMD MI355X (gfx950).

Drop-in entry points with the reference's signatures:
  * ``nanopyx_amd.eSRRF(...)`` — the workflow function (src/nanopyx/methods/esrrf/eSRRF_workflow.py:12)
  * ``nanopyx_amd.core.transform.eSRRF_ST`` / ``RadialGradientConvergence`` — Liquid Engine operators whose
    run type ``HIP_gfx950`` calls hand-written HIP kernels through the ctypes C ABI in include/npx_esrrf.h
  * ``calculate_eSRRF_temporal_correlations``, ``macro_pixel_corrector``
  * ``nanopyx_amd.SRRF(...)`` / ``core.transform.Radiality`` — SRRF radiality (src/nanopyx/methods/srrf/
    SRRF_workflow.py:9-76, src/nanopyx/core/transform/_le_radiality.pyx:27-363) on one fused kernel
  * ``nanopyx_amd.eSRRF3D(...)`` / ``core.transform.eSRRF3D_ST`` — eSRRF 3D (src/nanopyx/methods/esrrf_3d/
    eSRRF3D_workflow.py:8-119, src/nanopyx/core/transform/_le_esrrf3d.pyx:26-687)
  * ``run_esrrf_parameter_sweep`` / ``core.analysis.ParameterSweep`` — every (sensitivity, radius) reconstruction
    in one batched launch, scored by QnR on the host (src/nanopyx/core/analysis/parameter_sweep.py:11-139)
  * ``calculate_frc``, ``calculate_decorr_analysis``, ``calculate_error_map`` — the metrics the sweep scores with
"""
from .methods import (SRRF, eSRRF, eSRRF3D, Workflow, calculate_decorr_analysis, calculate_error_map, calculate_frc,
                      run_esrrf_parameter_sweep)
from .core.transform import (Radiality, RadialGradientConvergence, calculate_eSRRF_temporal_correlations, eSRRF_ST,
                             macro_pixel_corrector)

__version__ = "0.1.0"

__all__ = ["eSRRF", "eSRRF3D", "SRRF", "Workflow", "run_esrrf_parameter_sweep", "eSRRF_ST", "RadialGradientConvergence",
           "Radiality", "calculate_eSRRF_temporal_correlations", "macro_pixel_corrector", "calculate_frc",
           "calculate_decorr_analysis", "calculate_error_map"]
