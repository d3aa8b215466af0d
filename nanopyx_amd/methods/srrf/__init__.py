from .SRRF_workflow import SRRF

__all__ = ["SRRF"]
