from .esrrf import eSRRF, run_esrrf_parameter_sweep
from .esrrf_3d import eSRRF3D
from .srrf import SRRF
from .squirrel import calculate_decorr_analysis, calculate_error_map, calculate_frc
from .workflow import Workflow

__all__ = ["eSRRF", "eSRRF3D", "SRRF", "Workflow", "run_esrrf_parameter_sweep", "calculate_frc", "calculate_decorr_analysis",
           "calculate_error_map"]
