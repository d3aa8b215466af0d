"""eSRRF operator (``eSRRF_ST``) with the gfx950 run type.

Drop-in for NanoPyx's ``eSRRF`` Liquid Engine class (src/nanopyx/core/transform/_le_esrrf.pyx:21-296): same
constructor, same ``run`` / ``benchmark`` signature (:30-37), same (N, H*M, W*M) float32 output with a zero
border of ``int(radius*M)`` pixels. The run type ``HIP_gfx950`` replaces the CPU/OpenCL chain
CR xM -> Roberts -> CR x2M (twice) -> RGC (:172-195) by one fused kernel (libnpx_hip ``npx_esrrf_frames``).
``run_temporal`` is the fused eSRRF + temporal-correlation entry the workflow uses, so the per-frame maps never
leave the GPU.
"""
from __future__ import annotations

import numpy as np

from ... import _hip
from ...__liquid_engine__ import LiquidEngine
from .__interpolation_tools__ import check_image


def _is_tensor(x) -> bool:
    try:
        import torch
        return isinstance(x, torch.Tensor)
    except ImportError:  # pragma: no cover
        return False


def _esrrf_fht(image, M, radius, sensitivity, iw, pad_edges):
    """The use_fht branch (_le_esrrf.pyx:183-188): Fourier xM of the image, Roberts cross of the image, Fourier x2M
    of both gradients (npx_fht_interpolate, npx_roberts), then RGC on those arrays with pad_edges forwarded
    (npx_rgc); returns the (N, H*M, W*M) device tensor. The RGC reads the x2M gradients only at multiples of M
    (and the last row / column), so for M >= 2 only that lattice of them is computed into compact arrays
    (npx_fht_interpolate_lattice) and read by npx_rgc_lattice; radii that kernel does not cover take the full
    arrays and npx_rgc."""
    x = _hip._dev_stack(image)
    mag = _hip.fht_interpolate(x, M)
    gc, gr = _hip.roberts(x)
    if M >= 2:
        gci = _hip.fht_interpolate(gc, 2 * M, lattice=M)
        gri = _hip.fht_interpolate(gr, 2 * M, lattice=M)
        try:
            return _hip.rgc_lattice(gci, gri, mag, M, radius, sensitivity, iw, pad_edges)
        except _hip.NpxUnsupported:
            del gci, gri
    gci = _hip.fht_interpolate(gc, 2 * M)
    del gc
    gri = _hip.fht_interpolate(gr, 2 * M)
    del gr
    return _hip.rgc(gci, gri, mag, M, radius, sensitivity, iw, pad_edges)


class eSRRF(LiquidEngine):
    """eSRRF using the Liquid Engine and running as a single task."""

    def __init__(self, clear_benchmarks=False, testing=False, verbose=True):
        self._designation = "eSRRF_ST"
        super().__init__(clear_benchmarks=clear_benchmarks, testing=testing, verbose=verbose)

    def run(self, image, magnification: int = 5, radius: float = 1.5, sensitivity: float = 1, use_fht: bool = False,
            doIntensityWeighting: bool = True, pad_edges: bool = False, run_type=None):
        image = check_image(image, keep_uint16=True)
        return self._run(image, magnification=magnification, radius=radius, sensitivity=sensitivity,
                         use_fht=use_fht, doIntensityWeighting=doIntensityWeighting, pad_edges=pad_edges,
                         run_type=run_type)

    def benchmark(self, image, magnification: int = 5, radius: float = 1.5, sensitivity: float = 1,
                  use_fht: bool = False, doIntensityWeighting: bool = True, pad_edges: bool = False):
        image = check_image(image, keep_uint16=True)
        return super().benchmark(image, magnification=magnification, radius=radius, sensitivity=sensitivity,
                                 use_fht=use_fht, doIntensityWeighting=doIntensityWeighting, pad_edges=pad_edges)

    def _run_hip(self, image, magnification=5, radius=1.5, sensitivity=1, use_fht=False, doIntensityWeighting=True,
                 pad_edges=False):
        """
        @gpu
        """
        if use_fht:
            out = _esrrf_fht(image, int(magnification), radius, sensitivity, doIntensityWeighting, pad_edges)
            return out if _is_tensor(image) else out.cpu().numpy()
        # pad_edges is accepted and, as in the reference's non-FHT branch (_le_esrrf.pyx:194), not forwarded
        out = _hip.esrrf_frames(image, int(magnification), radius, sensitivity, doIntensityWeighting)
        return out if _is_tensor(image) else out.cpu().numpy()

    def run_temporal(self, image, magnification: int = 5, radius: float = 1.5, sensitivity: float = 1,
                     temporal_correlation="AVG", doIntensityWeighting: bool = True, run_type=None, stream=None):
        """eSRRF followed by calculate_eSRRF_temporal_correlations in one fused launch: returns the (H*M, W*M)
        map for one mode, or a {mode: map} dict for a sequence of modes, as device tensors."""
        image = check_image(image, keep_uint16=True)
        if run_type not in (None, "HIP_gfx950"):
            raise ValueError(f"{self._designation}: unknown run type {run_type!r}; available: ['HIP_gfx950']")
        modes = (temporal_correlation,) if isinstance(temporal_correlation, str) else tuple(temporal_correlation)
        import time
        t0 = time.perf_counter()
        outs = _hip.esrrf_temporal(image, int(magnification), radius, sensitivity, doIntensityWeighting, modes,
                                   stream=stream)
        self._last_runtype, self._last_time = "HIP_gfx950", time.perf_counter() - t0
        return outs[modes[0]] if isinstance(temporal_correlation, str) else outs
