from .eSRRF3D_workflow import eSRRF3D

__all__ = ["eSRRF3D"]
