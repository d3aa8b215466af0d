"""Image-quality metrics with the reference's package-level signatures (src/nanopyx/methods/squirrel/): the FRC/FIRE
resolution, the decorrelation resolution and the error map, as the eSRRF parameter sweep scores its maps."""
from .error_map import calculate_error_map
from .resolution import calculate_decorr_analysis, calculate_frc

__all__ = ["calculate_error_map", "calculate_frc", "calculate_decorr_analysis"]
