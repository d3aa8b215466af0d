"""``run_esrrf_parameter_sweep`` with the reference's signature and results
(src/nanopyx/methods/esrrf/parameter_sweep.py:6-81): the QnR map over sensitivities x radii (``return_qnr=True``)
or the (sensitivity, radius) pair that maximises it. The reconstructions come from batched GPU launches
(nanopyx_amd.core.analysis.ParameterSweep)."""
from __future__ import annotations

import numpy as np

from ...core.analysis.parameter_sweep import ParameterSweep


def run_esrrf_parameter_sweep(img: np.ndarray, magnification: int = 2, sensitivities: list = [1, 2],
                              radii: list = [1, 1.5], temporal_correlation: str = "AVG", use_decorr: bool = False,
                              plot_sweep=False, return_qnr=False, n_frames=None):
    ps = ParameterSweep()
    out = ps.run(img, magnification, sensitivity_array=sensitivities, radius_array=radii,
                 temporal_correlation=temporal_correlation, use_decorr=use_decorr, n_frames=n_frames)
    if plot_sweep:
        from matplotlib import pyplot as plt  # optional, as in the reference (:62-76)
        fig, ax = plt.subplots()
        ax.imshow(out, cmap="plasma")
        ax.set_xticks(np.arange(len(radii)), labels=radii)
        ax.set_yticks(np.arange(len(sensitivities)), labels=sensitivities)
        for i in range(len(sensitivities)):
            for j in range(len(radii)):
                ax.text(j, i, round(out[i, j], 2), ha="center", va="center", color="w")
        ax.set_title("Parameter Sweep QnR")
        ax.set_xlabel("Radii")
        ax.set_ylabel("Sensitivities")
        fig.tight_layout()
        plt.show()
    if return_qnr:
        return out
    sens_idx, rad_idx = np.unravel_index(np.argmax(out), out.shape)
    return sensitivities[sens_idx], radii[rad_idx]
