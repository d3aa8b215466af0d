from .parameter_sweep import ParameterSweep

__all__ = ["ParameterSweep"]
