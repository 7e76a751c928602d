"""Host-side mirror of Acts::KalmanFitter for the batched MI355X fit.

``KalmanFitter(detector, field, options).fit(events)`` plays the role of
calling ``Acts::KalmanFitter<Propagator<EigenStepper<>, DirectNavigator>>
::fit(begin, end, startParameters, kfOptions, surfaceSequence, tracks)``
(Core/include/Acts/TrackFitting/KalmanFitter.hpp:721-731) once per track,
for a whole batch: same options, same per-track error codes
(``status`` holds the reference enum values), same track-state content.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, fields

import numpy as np

from . import capi
from .detector import ConstantField, Detector, MUON_MASS
from .events import Events


@dataclass
class KalmanFitterOptions:
    """KalmanFitterOptions + PropagatorPlainOptions + StepperPlainOptions."""
    multiple_scattering: bool = True        # KalmanFitter.hpp:166
    energy_loss: bool = True                # KalmanFitter.hpp:169
    use_joseph: bool = False                # GainMatrixUpdater
    outlier_chi2_cut: float = 0.0           # SimpleOutlierFinder; <= 0 disables
    precision: int = capi.PRECISION_FP64
    smoother: int = capi.SMOOTHER_GAIN_MATRIX  # capi.SMOOTHER_MBF: as the Examples
    stepper: int = 0                        # capi.STEPPER_*
    step_tolerance: float = 1e-4            # StepperOptions.hpp:37
    step_size_cutoff: float = 0.0
    initial_step_size: float = 10000.0
    max_step_size: float = np.finfo(np.float64).max
    max_runge_kutta_step_trials: int = 10000
    max_steps: int = 1000                   # PropagatorOptions.hpp:30
    max_target_skipping: int = 100
    loop_protection: bool = True
    path_limit: float = np.finfo(np.float64).max
    loop_fraction: float = 0.5
    surface_tolerance: float = 1e-4
    near_limit: float = -0.1                # DirectNavigator.hpp:58 (-100 um)
    abs_pdg: int = 13
    mass: float = MUON_MASS
    abs_charge: float = 1.0
    # fitted parameters on a reference surface (KalmanFitterOptions::
    # referenceSurface): capi.REFERENCE_*, capi.EXTRAPOLATE_*, perigee point
    reference_surface: int = capi.REFERENCE_NONE
    reference_strategy: int = capi.EXTRAPOLATE_FIRST_OR_LAST
    reference_position: tuple = (0.0, 0.0, 0.0)
    # scheduling only; results do not depend on these
    length_sorted: int = 1                  # queue order: 0 input, 1 length, 2 direction, 3 length+momentum
    sched_inner_steps: int = 0              # 0 = library default
    sched_wait_64ths: int = 0               # 0 = library default
    sched_quarantine_steps: int = 0         # 0 = off
    kernel_timing: bool = False             # events around the kernels
    use_graph: bool = False                 # replay fits as a captured hipGraph
    sched_transport_skip_steps: int = 0     # 0 = off
    # capi.NAVIGATOR_DIRECT: each track's given surface sequence
    # (DirectNavigator); capi.NAVIGATOR_LAYERS: the surfaces found from the
    # geometry (Navigator), the sequence offsets giving each track's slots
    navigator: int = capi.NAVIGATOR_DIRECT
    # reverse filtering instead of smoothing (KalmanFitterOptions::
    # reverseFiltering, reverseFilteringCovarianceScaling; the Examples'
    # momentum threshold of SimpleReverseFilteringLogic)
    reverse_filtering: bool = False
    reserved_reverse: int = 0
    reverse_filtering_cov_scaling: float = 100.0
    reverse_filtering_momentum_threshold: float = 0.0

    def struct(self) -> capi.Options:
        o = capi.Options()
        for f in fields(self):
            v = getattr(self, f.name)
            if isinstance(v, (tuple, list)):
                v = (C.c_double * len(v))(*v)
            setattr(o, f.name, int(v) if isinstance(v, bool) else v)
        return o


STATE_ARRAYS = {
    # name: (dtype, per-state width)
    "state_surface": (np.int32, 1),
    "state_flags": (np.uint32, 1),
    "state_path": (np.float64, 1),
    "state_chi2": (np.float32, 1),
    "predicted": (np.float64, 6),
    "predicted_cov": (np.float64, 36),
    "filtered": (np.float64, 6),
    "filtered_cov": (np.float64, 36),
    "smoothed": (np.float64, 6),
    "smoothed_cov": (np.float64, 36),
    "jacobian": (np.float64, 36),
    "state_measurement": (np.int32, 1),
}
TRACK_ARRAYS = {
    "status": np.int32, "n_states": np.int32, "chi2": np.float32, "ndf": np.int32,
    "n_measurements": np.int32, "n_holes": np.int32, "n_outliers": np.int32,
    "n_steps": np.int32, "n_step_trials": np.int32, "reversed": np.int32,
}
# per-track parameter arrays, filled when the options ask for a reference
# surface: name -> width
TRACK_VECTOR_ARRAYS = {"reference_params": 6, "reference_cov": 36}
DEFAULT_STATE_OUTPUTS = ("state_surface", "state_flags", "state_path", "state_chi2",
                         "smoothed", "smoothed_cov")
ALL_STATE_OUTPUTS = tuple(STATE_ARRAYS)


class FitResult:
    """Per-track and per-state output arrays (numpy) of one batch."""

    def __init__(self, n_tracks: int, n_slots: int, outputs=DEFAULT_STATE_OUTPUTS):
        self.n_tracks = n_tracks
        for name, dt in TRACK_ARRAYS.items():
            setattr(self, name, np.zeros(n_tracks, dt))
        self.outputs = tuple(outputs)
        for name, w in TRACK_VECTOR_ARRAYS.items():
            setattr(self, name, np.zeros((n_tracks, w)) if name in self.outputs else None)
        for name in STATE_ARRAYS:
            dt, w = STATE_ARRAYS[name]
            if name in self.outputs:
                shape = (n_slots,) if w == 1 else (n_slots, w)
                setattr(self, name, np.zeros(shape, dt))
            else:
                setattr(self, name, None)

    def batch_out(self) -> capi.BatchOut:
        b = capi.BatchOut()
        for name in capi.BATCH_OUT_FIELDS:
            arr = getattr(self, name, None)
            setattr(b, name, arr.ctypes.data if arr is not None else None)
        return b


class KalmanFitter:
    """Batched Kalman fitter on one HIP device."""

    def __init__(self, detector: Detector, field: ConstantField,
                 options: KalmanFitterOptions | None = None, device: int = 0):
        self.lib = capi.load_library()
        self.detector = detector
        self.options = options or KalmanFitterOptions()
        self._surf = detector.surface_table()
        self._mats = detector.material_table()
        self._field = field.struct()
        self._opt = self.options.struct()
        ctx = C.c_void_p()
        capi.check(self.lib.acts_kf_create(self._surf, len(detector.surfaces), self._mats,
                                           len(detector.materials), C.byref(self._field),
                                           C.byref(self._opt), device, C.byref(ctx)),
                   "acts_kf_create")
        self.ctx = ctx

    def fit(self, events: Events, outputs=DEFAULT_STATE_OUTPUTS) -> FitResult:
        """Fit a batch held in host memory; blocks until the results are back."""
        res = FitResult(events.n_tracks, int(events.seq_offsets[-1]), outputs)
        bin_ = events.batch_in()
        bout = res.batch_out()
        capi.check(self.lib.acts_kf_fit_batch(self.ctx, C.byref(bin_), C.byref(bout), None),
                   "acts_kf_fit_batch")
        capi.check(self.lib.acts_kf_sync(self.ctx), "acts_kf_sync")
        return res

    def fit_device(self, bin_: capi.BatchIn, bout: capi.BatchOut, stream=None) -> None:
        """Fit a batch whose buffers are device memory; asynchronous on `stream`."""
        capi.check(self.lib.acts_kf_fit_batch_device(self.ctx, C.byref(bin_), C.byref(bout),
                                                     stream), "acts_kf_fit_batch_device")

    def track_states_device(self, bin_: capi.BatchIn, fitted: capi.BatchOut, truth_ptr,
                            out: capi.TrackStatesOut, stream=None) -> None:
        """Fill the track-state table of a fitted batch (device buffers);
        asynchronous on `stream`."""
        capi.check(self.lib.acts_kf_track_states_device(self.ctx, C.byref(bin_), C.byref(fitted),
                                                        truth_ptr, C.byref(out), stream),
                   "acts_kf_track_states_device")

    def sync(self) -> None:
        capi.check(self.lib.acts_kf_sync(self.ctx), "acts_kf_sync")

    def kernel_times(self) -> tuple[float, float]:
        """(forward, smoother) kernel milliseconds of the last fit."""
        f, s = C.c_double(), C.c_double()
        capi.check(self.lib.acts_kf_kernel_times(self.ctx, C.byref(f), C.byref(s)),
                   "acts_kf_kernel_times")
        return f.value, s.value

    def close(self) -> None:
        if getattr(self, "ctx", None):
            self.lib.acts_kf_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
