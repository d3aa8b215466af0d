"""Temporal correlations of eSRRF stacks on the GPU.

``calculate_eSRRF_temporal_correlations`` keeps the reference's signature and errors
(src/nanopyx/core/transform/sr_temporal_correlations.py:46-75): AVG = mean over frames, VAR = variance with
ddof 0, TAC2 = mean of lag-1 products of the mean-centred stack (calculate_tac2, :205-221). The moments are
computed in one streaming pass (libnpx_hip ``npx_temporal``).
"""
from __future__ import annotations

import numpy as np

from ... import _hip
from ._le_esrrf import _is_tensor


def calculate_eSRRF_temporal_correlations(im, correlation: str):
    if correlation not in ("AVG", "VAR", "TAC2"):
        raise ValueError("Type of correlation must be AVG, VAR or TAC2")
    out = _hip.temporal(im, (correlation,))[correlation]
    return out if _is_tensor(im) else out.cpu().numpy()


def calculate_tac2(rad_array):
    return calculate_eSRRF_temporal_correlations(rad_array, "TAC2")
