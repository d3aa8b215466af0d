"""Error map between a diffraction-limited reference and a super-resolved image (RSE / RSP).

Restates ``ErrorMap`` of src/nanopyx/core/transform/error_map.pyx:22-90 and its helpers ``calculate_alpha_beta`` /
``sigma_function_to_optimize`` (:124-146): the reference is Fourier-interpolated to the super-resolved grid, the
blur sigma that best maps the intensity-rescaled super-resolved image onto it is found with Brent's method on
(0, 4/2.35482 * magnification), and RSP is the Pearson correlation of the rescaled, blurred image with the
interpolated reference. The SciPy calls (gaussian_filter, linregress, brent) are the reference's own; host code.
"""
from __future__ import annotations

import numpy as np
from scipy.ndimage import gaussian_filter
from scipy.optimize import brent
from scipy.stats import linregress

from ..analysis.pearson_correlation import pearson_correlation
from .fht_interpolation import fht_space_interpolation


def calculate_alpha_beta(sigma: float, imRef: np.ndarray, imSR: np.ndarray) -> tuple:
    """Slope and intercept of the linear regression of imRef on the sigma-blurred imSR (:124-140)."""
    blurred = gaussian_filter(np.asarray(imSR), sigma)
    slope, intercept, r, p, se = linregress(np.asarray(blurred).ravel(), np.asarray(imRef).ravel())
    return slope, intercept


def sigma_function_to_optimize(sigma: float, imRef: np.ndarray, imSR: np.ndarray) -> float:
    """RMSE between imRef and the rescaled, sigma-blurred imSR (:142-146)."""
    alpha, beta = calculate_alpha_beta(sigma, imRef, imSR)
    scaled = gaussian_filter(np.asarray(imSR) * alpha + beta, sigma)
    return np.mean((scaled - np.asarray(imRef)) ** 2) ** 0.5


class ErrorMap:
    def __init__(self):
        self._vRSE = 0.0
        self._vRSP = 0.0
        self._alpha = 0.0
        self._beta = 0.0
        self._sigma = 0.0
        self.imRSE = None
        self.img_ref = None
        self.img_sr = None
        self.im_sr_intensity_scaled_blurred = None

    def optimise(self, img_ref: np.ndarray, img_sr: np.ndarray, fixedSigma=0) -> None:
        img_ref = np.asarray(img_ref).astype(np.float32)
        img_sr = np.asarray(img_sr).astype(np.float32)
        self.img_ref, self.img_sr = img_ref, img_sr
        magnification = int(img_sr.shape[0] / img_ref.shape[0])
        assert magnification == int(img_sr.shape[1] / img_ref.shape[1])
        if magnification > 1:
            img_ref_int = fht_space_interpolation(np.asarray([img_ref]), magnification)[0]
        else:
            img_ref_int = img_ref
        max_sigma_boundary = (4 / 2.35482) * magnification  # Nyquist-sampled reference (:61-63)
        sigma_linear = fixedSigma * magnification
        if fixedSigma == 0:
            sigma_linear = brent(sigma_function_to_optimize, args=(img_ref_int, img_sr), brack=(0, max_sigma_boundary),
                                 maxiter=1000)
        alpha, beta = calculate_alpha_beta(sigma_linear, img_ref_int, img_sr)
        # the reference keeps alpha, beta and sigma as float32 attributes and uses those values below
        self._alpha = float(np.float32(alpha))
        self._beta = float(np.float32(beta))
        self._sigma = float(np.float32(sigma_linear))
        self.im_sr_intensity_scaled_blurred = gaussian_filter(self._alpha * img_sr + self._beta, self._sigma).astype(
            np.float32)
        diff = self.im_sr_intensity_scaled_blurred - img_ref_int
        self.imRSE = np.abs(diff).astype(np.float32)
        self._vRSE = float(np.float32(np.mean(diff ** 2) ** 0.5))
        self._vRSP = float(np.float32(pearson_correlation(self.im_sr_intensity_scaled_blurred, img_ref_int)))

    def getRSE(self) -> float:
        return self._vRSE

    def getRSP(self) -> float:
        return self._vRSP

    def get_sigma(self) -> float:
        return self._sigma
