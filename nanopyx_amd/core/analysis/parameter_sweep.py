"""eSRRF parameter sweep on the GPU: every (sensitivity, radius) reconstruction of one stack from one batched launch.

Mirrors ``ParameterSweep`` of src/nanopyx/core/analysis/parameter_sweep.py:11-97. The reference runs one full
``eSRRF().run`` per (sensitivity, radius) pair (:33-37) and then ``calculate_eSRRF_temporal_correlations`` on its
frames (:40); here all pairs (and several magnifications) come out of one ``npx_esrrf_sweep`` launch: the
sensitivities share the ring sums and the radii of one magnification share the gradient lattice. AVG is computed
in-kernel; VAR and TAC2 fall back to one fused ``npx_esrrf_temporal`` call per pair (still on the GPU).

The reference's scoring of the reconstructions (ErrorMap RSP, FRC / decorrelation and the QnR map, :99-139) is not
part of this package: ``run`` raises NotImplementedError when it is requested, and ``reconstructions`` returns the
maps the scoring would consume.
"""
from __future__ import annotations

import numpy as np

from ... import _hip


class ParameterSweep:
    def __init__(self, doErrorMapping: bool = True, doFRCMapping: bool = True):
        self.doErrorMapping = doErrorMapping
        self.doFRCMapping = doFRCMapping

    def reconstructions(self, im, magnification, sensitivity_array, radius_array, temporal_correlation="AVG",
                        doIntensityWeighting=True, as_numpy=True):
        """Temporal reconstructions of every (sensitivity, radius): {M: array (n_sens, n_rad, H*M, W*M)} for each
        magnification in ``magnification`` (an int or a list); indices follow the reference's RSP_map[s, r]."""
        mags = [int(m) for m in (magnification if np.iterable(magnification) else [magnification])]
        sens = [float(s) for s in sensitivity_array]
        radii = [float(r) for r in radius_array]
        if temporal_correlation not in ("AVG", "VAR", "TAC2"):
            raise ValueError("Type of correlation must be AVG, VAR or TAC2")
        torch = _hip._torch()
        x = _hip._dev_stack(im)
        result = {}
        if temporal_correlation == "AVG":
            out = _hip.esrrf_sweep(x, mags, radii, sens, iw=doIntensityWeighting)
            for M in mags:
                maps = torch.stack([torch.stack([out.maps[(M, r, s)] for r in radii]) for s in sens])
                result[M] = maps
        else:
            for M in mags:
                rows = []
                for s in sens:
                    rows.append(torch.stack([
                        _hip.esrrf_temporal(x, M, r, s, doIntensityWeighting, (temporal_correlation,))[
                            temporal_correlation] for r in radii]))
                result[M] = torch.stack(rows)
        if as_numpy:
            torch.cuda.synchronize()
            result = {M: v.cpu().numpy() for M, v in result.items()}
        return result

    def run(self, im, magnification, sensitivity_array, radius_array, temporal_correlation="AVG",
            use_decorr=False, n_frames=None):
        """Reference signature (:17-26). Scoring (RSP / FRC / QnR) is not built: with doErrorMapping and
        doFRCMapping both False the reconstructions {M: (n_sens, n_rad, H*M, W*M)} are returned instead."""
        if self.doErrorMapping or self.doFRCMapping or use_decorr:
            raise NotImplementedError(
                "sweep scoring (ErrorMap RSP, FRC / decorrelation, QnR) is not part of nanopyx_amd; use "
                "ParameterSweep(doErrorMapping=False, doFRCMapping=False).run(...) or .reconstructions(...) for the "
                "reconstructions it would score")
        if n_frames is not None:
            raise NotImplementedError("n_frames slicing feeds only the RSP scoring, which is not built")
        return self.reconstructions(im, magnification, sensitivity_array, radius_array, temporal_correlation)
