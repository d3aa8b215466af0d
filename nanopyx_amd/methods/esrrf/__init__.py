from .eSRRF_workflow import eSRRF
from .parameter_sweep import run_esrrf_parameter_sweep

__all__ = ["eSRRF", "run_esrrf_parameter_sweep"]
