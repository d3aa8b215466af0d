"""``eSRRF()`` workflow function (src/nanopyx/methods/esrrf/eSRRF_workflow.py:12-121) on the gfx950 path.

Same signature and output: frames are split into timepoints of ``frames_per_timepoint`` (0 = all), each
timepoint becomes one (H*M, W*M) map of the requested temporal correlation, optionally macro-pixel corrected,
and the (T, H*M, W*M) float32 result is squeezed. On this path each timepoint is one fused launch
(eSRRF + AVG/VAR/TAC2, ``npx_esrrf_temporal``) followed by the macro-pixel kernel; the stack is uploaded once and
the per-frame maps never leave the GPU (radii with int(2 R) > 6 run the reference's separate steps on the GPU and
reduce their per-frame maps). A host (NumPy) stack of 64 MiB or more is streamed instead
(``nanopyx_amd.upload``): pieces go to the device through pinned staging buffers while a serving session's jobs, one
per timepoint, already run on the frames that have arrived. With ``use_fht`` each timepoint runs the
Fourier-interpolation branch of the operator (per-frame maps on the GPU), then ``npx_temporal`` and the macro-pixel
kernel.
"""
from __future__ import annotations

import numpy as np

from ...__agent__ import Agent
from ...core.transform import eSRRF_ST
from ...core.transform.__interpolation_tools__ import check_image
from ... import _hip, upload


def eSRRF(image, magnification: int = 5, radius: float = 1.5, sensitivity: float = 1, frames_per_timepoint: int = 0,
          temporal_correlation: str = "AVG", use_fht: bool = False, doIntensityWeighting: bool = True,
          macro_pixel_correction: bool = True, pad_edges: bool = False, _force_run_type=None):
    """Perform eSRRF analysis on an (N, H, W) stack; returns the squeezed (T, H*M, W*M) float32 result."""
    if temporal_correlation not in ("AVG", "VAR", "TAC2"):
        raise ValueError("Type of correlation must be AVG, VAR or TAC2")
    import torch

    is_tensor = isinstance(image, torch.Tensor)
    if not is_tensor and not use_fht and not _hip._wide(radius):
        x = check_image(image, keep_uint16=True)
        if x.nbytes >= upload.STREAM_MIN_BYTES:
            # a large host stack: uploaded in pieces while the timepoints' kernels already run on the arrived frames
            if _force_run_type not in (None, "HIP_gfx950"):
                raise ValueError(f"eSRRF: unknown run type {_force_run_type!r}; available: ['HIP_gfx950']")
            fpt = x.shape[0] if frames_per_timepoint in (0, None) or frames_per_timepoint > x.shape[0] \
                else frames_per_timepoint
            out = upload.esrrf_temporal_streamed(np.ascontiguousarray(x), magnification, radius, sensitivity,
                                                 doIntensityWeighting, temporal_correlation, fpt)
            if macro_pixel_correction:
                out = _hip.macro_pixel_(out, int(magnification))
            return torch.squeeze(out).cpu().numpy().astype(np.float32)
    N = image.shape[0]
    if frames_per_timepoint == 0:
        frames_per_timepoint = N
    elif frames_per_timepoint > N:
        frames_per_timepoint = N
    T = N // frames_per_timepoint + (1 if N % frames_per_timepoint else 0)
    M = int(magnification)
    stack = _hip._dev_stack(check_image(image, keep_uint16=True))
    out = torch.zeros((T, image.shape[1] * M, image.shape[2] * M), dtype=torch.float32, device=stack.device)
    op = eSRRF_ST(verbose=False)
    for i in range(T):
        part = stack[frames_per_timepoint * i:frames_per_timepoint * (i + 1)]
        run_type = _force_run_type or Agent.get_run_type(op, (part,), {})
        if use_fht:
            # Fourier interpolation branch: per-frame maps (eSRRF_ST.run with use_fht, pad_edges forwarded), then
            # the temporal correlation of the maps
            maps = op.run(part, magnification=M, radius=radius, sensitivity=sensitivity, use_fht=True,
                          doIntensityWeighting=doIntensityWeighting, pad_edges=pad_edges, run_type=run_type)
            m = _hip.temporal(maps, (temporal_correlation,))[temporal_correlation]
        else:
            m = op.run_temporal(part, magnification=M, radius=radius, sensitivity=sensitivity,
                                temporal_correlation=temporal_correlation, doIntensityWeighting=doIntensityWeighting,
                                run_type=run_type)
        Agent._inform(op)
        if macro_pixel_correction:
            m = _hip.macro_pixel_(m.unsqueeze(0).contiguous(), M)[0]
        out[i] = m
    out = torch.squeeze(out)
    return out if is_tensor else out.cpu().numpy().astype(np.float32)
