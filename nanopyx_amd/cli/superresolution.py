"""eSRRF from the command line: ``python -m nanopyx_amd.cli.superresolution INPUT.tif OUTPUT.tif [options]``.

Mirrors ``linkinpy_eSRRF`` / ``linkinpy_eSRRF_main`` of src/nanopyx/cli/superresolution.py:76-145 (same arguments,
defaults and spellings), running the workflow ``nanopyx_amd.eSRRF`` on the GPU. TIFF files are read and written
with nanopyx_amd.io (uncompressed baseline TIFF; see that module).
"""
from __future__ import annotations

import argparse

from ..io import imread, imwrite


def linkinpy_eSRRF(input_image: str, output_image: str, magnification: int = 5, radius: float = 1.5,
                   sensitivity: float = 1.0, frames_per_timepoint: int = 0, temporal_correlation: str = "AVG",
                   use_fht: bool = False, doIntensityWeighting: bool = True, macro_pixel_correction: bool = True,
                   pad_edges: bool = False) -> None:
    from ..methods import eSRRF
    image = imread(input_image)
    output = eSRRF(image=image, magnification=magnification, radius=radius, sensitivity=sensitivity,
                   frames_per_timepoint=frames_per_timepoint, temporal_correlation=temporal_correlation,
                   use_fht=use_fht, doIntensityWeighting=doIntensityWeighting,
                   macro_pixel_correction=macro_pixel_correction, pad_edges=pad_edges)
    imwrite(output_image, output)


def parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description="Run NanoPyx eSRRF (HIP_gfx950).")
    p.add_argument("input_image", help="Input image path.")
    p.add_argument("output_image", help="Output image path.")
    p.add_argument("--magnification", type=int, default=5)
    p.add_argument("--radius", type=float, default=1.5)
    p.add_argument("--sensitivity", type=float, default=1.0)
    p.add_argument("--frames-per-timepoint", "--frames_per_timepoint", dest="frames_per_timepoint", type=int,
                   default=0)
    p.add_argument("--temporal-correlation", "--temporal_correlation", dest="temporal_correlation", default="AVG")
    p.add_argument("--use-fht", "--use_fht", dest="use_fht", action=argparse.BooleanOptionalAction, default=False)
    p.add_argument("--intensity-weighting", "--doIntensityWeighting", dest="doIntensityWeighting",
                   action=argparse.BooleanOptionalAction, default=True)
    p.add_argument("--macro-pixel-correction", "--macro_pixel_correction", dest="macro_pixel_correction",
                   action=argparse.BooleanOptionalAction, default=True)
    p.add_argument("--pad-edges", "--pad_edges", dest="pad_edges", action=argparse.BooleanOptionalAction,
                   default=False)
    return p


def linkinpy_eSRRF_main(argv=None) -> None:
    args = parser().parse_args(argv)
    linkinpy_eSRRF(input_image=args.input_image, output_image=args.output_image, magnification=args.magnification,
                   radius=args.radius, sensitivity=args.sensitivity, frames_per_timepoint=args.frames_per_timepoint,
                   temporal_correlation=args.temporal_correlation, use_fht=args.use_fht,
                   doIntensityWeighting=args.doIntensityWeighting,
                   macro_pixel_correction=args.macro_pixel_correction, pad_edges=args.pad_edges)


if __name__ == "__main__":
    linkinpy_eSRRF_main()
