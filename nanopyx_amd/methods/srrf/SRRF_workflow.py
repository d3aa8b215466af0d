"""``SRRF()`` workflow function (src/nanopyx/methods/srrf/SRRF_workflow.py:9-76) on the gfx950 path.

Same signature and output: the (N, H*M, W*M) float32 radiality maps of every frame, weighted by the xM
Catmull-Rom magnification of the frame, macro-pixel corrected per frame by default. The reference runs
CRShiftAndMagnify and then Radiality; here both are one launch (``npx_radiality`` magnifies the intensity inside
the kernel) followed by the macro-pixel kernel over all frames, with the stack uploaded once.
"""
from __future__ import annotations

import numpy as np

from ... import _hip
from ...__agent__ import Agent
from ...core.transform.__interpolation_tools__ import check_image
from ...core.transform._le_radiality import Radiality


def SRRF(image, magnification=5, ringRadius=0.5, border=0, radialityPositivityConstraint=True,
         doIntensityWeighting=True, macro_pixel_correction=True, _force_run_type=None):
    """Perform SRRF analysis on an (N, H, W) stack; returns the (N, H*M, W*M) float32 radiality maps."""
    import torch

    is_tensor = isinstance(image, torch.Tensor)
    M = int(magnification)
    stack = _hip._dev_stack(check_image(image, keep_uint16=True))
    op = Radiality(verbose=False)
    run_type = _force_run_type or Agent.get_run_type(op, (stack,), {})
    if run_type not in (None, "HIP_gfx950"):
        raise ValueError(f"Radiality: unknown run type {run_type!r}; available: ['HIP_gfx950']")
    out = op.run_fused(stack, magnification=M, ringRadius=ringRadius, border=border,
                       radialityPositivityConstraint=radialityPositivityConstraint,
                       doIntensityWeighting=doIntensityWeighting)
    Agent._inform(op)
    if macro_pixel_correction:
        out = _hip.macro_pixel_(out, M)
    return out if is_tensor else out.cpu().numpy().astype(np.float32)
