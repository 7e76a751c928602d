"""acts_amd -- MI355X-native batched Kalman track fitter.

Drop-in for the reference path Acts::KalmanFitter::fit with
Propagator<EigenStepper<>, DirectNavigator> (or the Navigator, which finds
the surfaces from the geometry), GainMatrixUpdater and GainMatrixSmoother,
run as hand-written HIP kernels behind the C ABI of include/acts_kf.h (one
lane per track, persistent wavefronts fed from a track queue).
"""
from . import capi  # noqa: F401
from .detector import (ConstantField, Detector, FieldMap, barrel_detector,  # noqa: F401
                       barrel_endcap_detector, double_layer_detector)
from .events import (Events, generate, gen_config, with_state_slots, workload,  # noqa: F401
                     workload_options)
from .fitter import FitResult, KalmanFitter, KalmanFitterOptions  # noqa: F401

__all__ = [
    "capi", "ConstantField", "Detector", "FieldMap", "barrel_detector", "barrel_endcap_detector",
    "double_layer_detector", "Events", "generate", "gen_config", "workload",
    "workload_options", "with_state_slots", "FitResult", "KalmanFitter",
    "KalmanFitterOptions",
]
