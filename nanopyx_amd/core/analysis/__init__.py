from .decorr import DecorrAnalysis
from .frc import FIRECalculator
from .parameter_sweep import ParameterSweep

__all__ = ["ParameterSweep", "FIRECalculator", "DecorrAnalysis"]
