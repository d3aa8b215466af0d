"""``eSRRF3D()`` workflow function (src/nanopyx/methods/esrrf_3d/eSRRF3D_workflow.py:8-119) on the gfx950 path.

Same signature and output: the (F, S, R, C) stack is split into timepoints of ``frames_per_timepoint`` frames (0 =
all); each becomes one (S*Mz, R*Mxy, C*Mxy) volume of the eSRRF 3D operator (running mean or std over its frames),
macro-pixel corrected slice by slice at magnification_xy by default; returns the (T, S*Mz, R*Mxy, C*Mxy) float32
array. The stack is uploaded once; each timepoint is one npx_esrrf3d call and one macro-pixel launch.
"""
from __future__ import annotations

import numpy as np

from ... import _hip
from ...__agent__ import Agent
from ...core.transform._le_esrrf3d import eSRRF3D as eSRRF3D_ST


def eSRRF3D(img, magnification_xy=2, magnification_z=2, radius: float = 1.5, PSF_ratio: float = 2.8,
            voxel_ratio: float = 4.0, sensitivity: float = 1, frames_per_timepoint: int = 0, mode: str = "average",
            doIntensityWeighting: bool = True, macro_pixel_correction: bool = True, _force_run_type=None):
    """Perform eSRRF 3D analysis on an (F, S, R, C) stack; returns (T, S*Mz, R*Mxy, C*Mxy) float32."""
    import torch

    is_tensor = isinstance(img, torch.Tensor)
    F = img.shape[0]
    if frames_per_timepoint == 0 or frames_per_timepoint > F:
        frames_per_timepoint = F
    T = F // frames_per_timepoint + (1 if F % frames_per_timepoint else 0)
    Mxy, Mz = int(magnification_xy), int(magnification_z)
    x = img if is_tensor else np.asarray(img)
    if not is_tensor and x.dtype not in (np.float32, np.uint16):
        x = x.astype(np.float32)
    x = (x if is_tensor else torch.from_numpy(np.ascontiguousarray(x))).cuda()
    out = torch.zeros((T, img.shape[1] * Mz, img.shape[2] * Mxy, img.shape[3] * Mxy), dtype=torch.float32,
                      device=x.device)
    op = eSRRF3D_ST(verbose=False)
    for i in range(T):
        part = x[frames_per_timepoint * i:frames_per_timepoint * (i + 1)]
        run_type = _force_run_type or Agent.get_run_type(op, (part,), {})
        v = op.run(part, magnification_xy=Mxy, magnification_z=Mz, radius=radius, PSF_ratio=PSF_ratio,
                   voxel_ratio=voxel_ratio, sensitivity=sensitivity, mode=mode,
                   doIntensityWeighting=doIntensityWeighting, run_type=run_type)
        Agent._inform(op)
        if v is None:
            raise ValueError("eSRRF3D: the radius does not fit the image (see the warning above)")
        if macro_pixel_correction:
            v = _hip.macro_pixel_(v.contiguous(), Mxy)
        out[i] = v
    return out if is_tensor else out.cpu().numpy().astype(np.float32)
