from .esrrf import eSRRF, run_esrrf_parameter_sweep
from .workflow import Workflow

__all__ = ["eSRRF", "Workflow", "run_esrrf_parameter_sweep"]
