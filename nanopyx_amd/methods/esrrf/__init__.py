from .eSRRF_workflow import eSRRF

__all__ = ["eSRRF"]
