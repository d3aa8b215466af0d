"""eSRRF 3D operator with the gfx950 run type.

Drop-in for src/nanopyx/core/transform/_le_esrrf3d.pyx:26-687: same constructor, ``run(image, magnification_xy=2,
magnification_z=2, radius=1.5, PSF_ratio=2.8, voxel_ratio=4.0, sensitivity=1, mode="average",
doIntensityWeighting=True, run_type=None)`` with the same checks (:38-56: a 3D image becomes one frame, anything not
4D, a radius above half the rows or a quarter of the columns, or radius_z = radius * PSF_ratio / voxel_ratio above
half the slices prints a warning and returns None), returning the (S*Mz, R*Mxy, C*Mxy) float32 volume. The
HIP_gfx950 run type is libnpx_hip ``npx_esrrf3d`` over the whole stack (per frame chunk: gradients, xMxy
Catmull-Rom of intensity and gradients, and the per-voxel convergence folded into the running mean or std).
"""
from __future__ import annotations

import numpy as np

from ... import _hip
from ...__liquid_engine__ import LiquidEngine
from ._le_esrrf import _is_tensor


class eSRRF3D(LiquidEngine):
    """eSRRF 3D using the Liquid Engine."""

    def __init__(self, clear_benchmarks=False, testing=False, verbose=True):
        self._designation = "eSRRF_3D"
        super().__init__(clear_benchmarks=clear_benchmarks, testing=testing, verbose=verbose)

    @staticmethod
    def _checked(image, radius, PSF_ratio, voxel_ratio):
        radius_z = radius * PSF_ratio / voxel_ratio
        if len(image.shape) == 3:
            image = image.reshape((1, image.shape[0], image.shape[1], image.shape[2]))
        if len(image.shape) != 4:
            print("Warning:image must either be 3D or 4D. If 3D, it will be reshaped to 4D.")
            return None, radius_z
        if radius > image.shape[2] / 2 or radius * 2 > image.shape[3] / 2:
            print("Warning: Radius is too big for the image. The radius must be smaller than half the number of "
                  "columns and half number of rows of the image.")
            return None, radius_z
        if radius_z > image.shape[1] / 2:
            print("Warning: Radius_z is too big for the image. The radius_z must be smaller than half of number of Z "
                  "planes. Radius_z is automatically calculated from radius * PSF_ratio / voxel_ratio.")
            return None, radius_z
        if not _is_tensor(image) and image.dtype not in (np.float32, np.uint16):
            image = image.astype(np.float32)
        return image, radius_z

    def run(self, image, magnification_xy: int = 2, magnification_z: int = 2, radius: float = 1.5,
            PSF_ratio: float = 2.8, voxel_ratio: float = 4.0, sensitivity: float = 1, mode: str = "average",
            doIntensityWeighting: bool = True, run_type=None):
        image, radius_z = self._checked(image, radius, PSF_ratio, voxel_ratio)
        if image is None:
            return None
        return self._run(image, magnification_xy=magnification_xy, magnification_z=magnification_z, radius=radius,
                         radius_z=radius_z, voxel_ratio=voxel_ratio, sensitivity=sensitivity, mode=mode,
                         doIntensityWeighting=doIntensityWeighting, run_type=run_type)

    def benchmark(self, image, magnification_xy: int = 5, magnification_z: int = 5, radius: float = 1.5,
                  PSF_ratio: float = 2.8, voxel_ratio: float = 4.0, sensitivity: float = 1, mode: str = "average",
                  doIntensityWeighting: bool = True):
        image, radius_z = self._checked(image, radius, PSF_ratio, voxel_ratio)
        if image is None:
            return None
        return super().benchmark(image, magnification_xy=magnification_xy, magnification_z=magnification_z,
                                 radius=radius, radius_z=radius_z, voxel_ratio=voxel_ratio, sensitivity=sensitivity,
                                 mode=mode, doIntensityWeighting=doIntensityWeighting)

    def _run_hip(self, image, magnification_xy=5, magnification_z=5, radius=1.5, radius_z=1.5, voxel_ratio=4.0,
                 sensitivity=1, mode="average", doIntensityWeighting=True):
        """
        @gpu
        """
        out = _hip.esrrf3d(image, int(magnification_xy), int(magnification_z), radius, radius_z, sensitivity, mode,
                           doIntensityWeighting)
        return out if _is_tensor(image) else out.cpu().numpy()
