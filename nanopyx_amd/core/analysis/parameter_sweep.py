"""eSRRF parameter sweep: every (sensitivity, radius) reconstruction of one stack from one batched GPU launch, scored
by the reference's QnR.

Mirrors ``ParameterSweep`` of src/nanopyx/core/analysis/parameter_sweep.py:11-139. The reference runs one full
``eSRRF().run`` per (sensitivity, radius) pair (:33-37) and then ``calculate_eSRRF_temporal_correlations`` on its
frames (:40, :47-50); here the reconstructions of all pairs come out of ``npx_esrrf_sweep`` launches (the
sensitivities share the ring sums, the radii of one magnification share the gradient lattice): one over all frames
(for RSP), one over the odd and one over the even frames (for FRC), one per ``n_frames`` slice when slicing. AVG is
computed in-kernel; VAR and TAC2 use one fused ``npx_esrrf_temporal`` call per pair (still on the GPU).

Scoring is the reference's, restated on the host as the reference runs it: RSP from ``ErrorMap`` (:99-102), the FIRE
number of the odd/even reconstructions (:104-107), and QnR = 2 RSP nFRC / (RSP + nFRC) with nFRC the logistic
rescaling of the FRC map (:109-139); with ``use_decorr=True`` the decorrelation resolution of the all-frames
reconstruction replaces the FIRE number (:43-46).
"""
from __future__ import annotations

import numpy as np

from ... import _hip
from ..transform.error_map import ErrorMap
from .decorr import DecorrAnalysis
from .frc import FIRECalculator


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
                result[M] = torch.stack([torch.stack([out.maps[(M, r, s)] for r in radii]) for s in sens])
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

    def run(self, im: np.ndarray, magnification: int, sensitivity_array: list, radius_array: list,
            temporal_correlation: str = "AVG", use_decorr: bool = False, n_frames=None):
        """QnR map (n_sens, n_rad) of the reference's sweep (:17-97).

        With ``n_frames`` the reference re-creates its (slices, sensitivities, radii) RSP array inside the loop over
        pairs and replaces it by its mean over slices (:55-65), so after the loop only the last pair's sliced RSP is
        non-zero and every other pair scores 0; that result is reproduced here."""
        im = np.asarray(im)
        if im.ndim == 2:
            im = im[None]
        M = int(magnification)
        sens, radii = list(sensitivity_array), list(radius_array)
        ns, nr = len(sens), len(radii)
        rsp = np.zeros((ns, nr))
        frc = np.zeros((ns, nr))
        x = _hip._dev_stack(im)

        def rec(stack, s=sens, r=radii):
            return self.reconstructions(stack, M, s, r, temporal_correlation)[M]

        full = None
        if (self.doErrorMapping and n_frames is None) or (self.doFRCMapping and use_decorr):
            full = rec(x)
        if self.doErrorMapping:
            if n_frames is None:
                for si in range(ns):
                    for ri in range(nr):
                        rsp[si, ri] = self.calculate_rsp(im, full[si, ri])
            else:
                n_slices = im.shape[0] // n_frames + (1 if im.shape[0] % n_frames else 0)
                per = [self.calculate_rsp(im[i * n_frames:(i + 1) * n_frames],
                                          rec(x[i * n_frames:(i + 1) * n_frames], [sens[-1]], [radii[-1]])[0, 0])
                       for i in range(n_slices)]
                rsp[ns - 1, nr - 1] = np.mean(per)
        if self.doFRCMapping:
            if use_decorr:
                for si in range(ns):
                    for ri in range(nr):
                        decorr = DecorrAnalysis()
                        decorr.run_analysis(full[si, ri])
                        frc[si, ri] = decorr.resolution
            else:
                odd = rec(x[1::2].contiguous())
                even = rec(x[::2].contiguous())
                for si in range(ns):
                    for ri in range(nr):
                        frc[si, ri] = self.calculate_frc(odd[si, ri], even[si, ri])
        return self.calculate_qnr_score(rsp, frc)

    def calculate_rsp(self, im, reconstruction):
        """RSP of a reconstruction against the temporal mean of the stack `im` (:99-102)."""
        error_map = ErrorMap()
        error_map.optimise(np.mean(np.asarray(im), axis=0), np.asarray(reconstruction))
        return error_map.getRSP()

    def calculate_frc(self, im_odd, im_even):
        return FIRECalculator().calculate_fire_number(np.asarray(im_odd), np.asarray(im_even))

    def logistic_image_conversion(self, im, min_val=None, max_val=None):
        """Maps [min, max] of `im` logistically onto (0.075, 0.925) (:113-133)."""
        im = np.asarray(im)
        if min_val is None:
            min_val = np.min(im)
        if max_val is None:
            max_val = np.max(im)
        M1, M2 = 0.075, 0.925
        A1 = np.log((1 - M1) / M1)
        A2 = np.log((1 - M2) / M2)
        x0 = (A2 * max_val - A1 * min_val) / (A2 - A1)
        k = 1 / (x0 - max_val) * A1
        return np.asarray([1 / (np.exp(-k * (image - x0)) + 1) for image in im])

    def calculate_qnr_score(self, RSP: np.ndarray, FRC: np.ndarray):
        assert RSP.shape == FRC.shape
        nFRC = self.logistic_image_conversion(FRC)
        return (2 * np.asarray(RSP) * np.asarray(nFRC)) / (np.asarray(RSP) + np.asarray(nFRC))
