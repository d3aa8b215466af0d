from ._le_esrrf import eSRRF
from ._le_esrrf import eSRRF as eSRRF_ST
from ._le_esrrf3d import eSRRF3D as eSRRF3D_ST
from ._le_radial_gradient_convergence import RadialGradientConvergence
from ._le_interpolation_catmull_rom import ShiftAndMagnify as CRShiftAndMagnify
from ._le_radiality import Radiality
from ._le_roberts_cross_gradients import GradientRobertsCross
from .interpolation import fht_space_interpolation
from .mpcorrector import macro_pixel_corrector
from .sr_temporal_correlations import calculate_eSRRF_temporal_correlations

__all__ = ["eSRRF_ST", "eSRRF", "eSRRF3D_ST", "RadialGradientConvergence", "Radiality", "CRShiftAndMagnify", "GradientRobertsCross",
           "fht_space_interpolation", "macro_pixel_corrector", "calculate_eSRRF_temporal_correlations"]
