"""``run_esrrf_parameter_sweep`` with the reference's signature (src/nanopyx/methods/esrrf/parameter_sweep.py:6-16).

The reference returns the QnR map (or its arg-max); QnR scoring is not part of nanopyx_amd, so with
``return_qnr=False`` and no scoring this returns the reconstructions of every (sensitivity, radius) instead, computed in
one batched GPU launch (nanopyx_amd.core.analysis.ParameterSweep)."""
from __future__ import annotations

from ...core.analysis.parameter_sweep import ParameterSweep


def run_esrrf_parameter_sweep(img, magnification: int = 2, sensitivities: list = [1, 2], radii: list = [1, 1.5],
                              temporal_correlation: str = "AVG", use_decorr: bool = False, plot_sweep=False,
                              return_qnr=False, n_frames=None, score=True):
    if plot_sweep:
        raise NotImplementedError("plotting the QnR map needs the sweep scoring, which is not built")
    ps = ParameterSweep(doErrorMapping=score, doFRCMapping=score)
    out = ps.run(img, magnification, sensitivity_array=sensitivities, radius_array=radii,
                 temporal_correlation=temporal_correlation, use_decorr=use_decorr, n_frames=n_frames)
    return out[int(magnification)] if not isinstance(magnification, (list, tuple)) else out
