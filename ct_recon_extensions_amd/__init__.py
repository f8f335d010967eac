"""MI355X-native (gfx950) 2-D CT reconstruction hot path.

Operators (torch, NHWC, float32/float64, autograd through hand-written
transpose kernels) mirror the reference's python/tf_ctrecon.py; the C ABI is
include/ctr_api.h, the reference's numpy C ABI is include/numpy_ctext.h and
its C++ class API is include/bp_fan_2d.h and siblings.
"""
from .functional import (fan_backprojection_2d, fan_backprojection_2d_angle,  # noqa: F401
                         fan_backprojection_2d_angle_grad, fan_backprojection_2d_grad,
                         fan_filter, fan_weighting, fan_weighting_grad,
                         helical_interpolation,
                         parallel_backprojection_2d, parallel_backprojection_2d_grad,
                         parallel_single_view_bp_2d,
                         parallel_projection_2d, parallel_projection_2d_grad, ramp_filter,
                         ramp_filter_grad)
from .torch_ops import bppar2d_forward, bppar2d_prep_forward  # noqa: F401
from .geometry import (FanBackprojection2DAngleParam, FanBackprojection2DParam,  # noqa: F401
                       FanWeightingParam, FilterParam, HelicalInterpolationParam,
                       ParallelBackprojection2DParam, ParallelProjection2DParam)

__version__ = "0.3.0"
