"""acts_amd -- MI355X-native batched Kalman track fitter.

Drop-in for the reference path Acts::KalmanFitter::fit with
Propagator<EigenStepper<>, DirectNavigator>, GainMatrixUpdater and
GainMatrixSmoother, run as hand-written HIP kernels (one lane per track)
behind the C ABI of include/acts_kf.h (one lane per track, persistent
wavefronts fed from a track queue).
"""
from . import capi  # noqa: F401
from .detector import (ConstantField, Detector, FieldMap, barrel_detector,  # noqa: F401
                       barrel_endcap_detector, double_layer_detector)
from .events import Events, generate, gen_config, workload, workload_options  # noqa: F401
from .fitter import FitResult, KalmanFitter, KalmanFitterOptions  # noqa: F401

__all__ = [
    "capi", "ConstantField", "Detector", "FieldMap", "barrel_detector", "barrel_endcap_detector",
    "double_layer_detector", "Events", "generate", "gen_config", "workload",
    "workload_options", "FitResult", "KalmanFitter",
    "KalmanFitterOptions",
]
