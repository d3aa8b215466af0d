from .esrrf import eSRRF
from .workflow import Workflow

__all__ = ["eSRRF", "Workflow"]
