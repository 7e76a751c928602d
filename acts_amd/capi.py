"""ctypes mirror of include/acts_kf.h and include/acts_kf_gen.h.

The native library is loaded from the package directory; it must have been
built beforehand (``python -m acts_amd.build``).  Nothing here falls back to
a CPU implementation: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# ACTS_KF_LIBRARY selects another build of the same library (e.g. a variant
# compiled with different flags); default: the in-tree build.
LIB_PATH = os.environ.get("ACTS_KF_LIBRARY", os.path.join(PKG_DIR, "libacts_kf.so"))

# ---- constants (acts_kf.h) -------------------------------------------------
OK = 0
ERR_INVALID_ARGUMENT = -1
ERR_DEVICE = -2
ERR_OUT_OF_MEMORY = -3
ERR_UNSUPPORTED = -4

CAT_KALMAN, CAT_PROPAGATOR, CAT_STEPPER, CAT_SURFACE, CAT_FIELD, CAT_EXTRAPOLATION = (
    0x100, 0x200, 0x300, 0x400, 0x500, 0x600)
CAT_LIBRARY = 0x700
STATUS_OK = 0
STATUS_UPDATE_FAILED = CAT_KALMAN | 1
STATUS_SMOOTH_FAILED = CAT_KALMAN | 2
STATUS_NO_MEASUREMENT_FOUND = CAT_KALMAN | 4
STATUS_REVERSE_PROPAGATION_FAILED = CAT_KALMAN | 5
STATUS_INCONSISTENT_TRACK_STATES = CAT_KALMAN | 6
STATUS_PROPAGATION_FAILURE = CAT_PROPAGATOR | 1
STATUS_STEP_COUNT_LIMIT = CAT_PROPAGATOR | 2
STATUS_NEXT_TARGET_LIMIT = CAT_PROPAGATOR | 3
STATUS_STEP_SIZE_STALLED = CAT_STEPPER | 1
STATUS_STEP_INVALID = CAT_STEPPER | 2
STATUS_STEP_ADJUSTMENT_FAILED = CAT_STEPPER | 3
STATUS_NOT_ON_SURFACE = CAT_SURFACE | 1
STATUS_FIELD_OUT_OF_BOUNDS = CAT_FIELD | 1
STATUS_NO_COMPATIBLE_STATE = CAT_EXTRAPOLATION | 1
STATUS_REFERENCE_UNREACHABLE = CAT_EXTRAPOLATION | 2
STATUS_STATE_CAPACITY = CAT_LIBRARY | 1

STATE_HAS_MEASUREMENT = 1 << 0
STATE_HAS_PARAMETERS = 1 << 1
STATE_IS_OUTLIER = 1 << 2
STATE_IS_HOLE = 1 << 3
STATE_HAS_MATERIAL = 1 << 4

SURFACE_PLANE, SURFACE_CYLINDER, SURFACE_DISC, SURFACE_LINE = 0, 1, 2, 3
REFERENCE_NONE, REFERENCE_PERIGEE = 0, 1
EXTRAPOLATE_FIRST, EXTRAPOLATE_LAST, EXTRAPOLATE_FIRST_OR_LAST = 0, 1, 2
SURFACE_SENSITIVE = 1
NAVIGATOR_DIRECT, NAVIGATOR_LAYERS = 0, 1

FIELD_CONSTANT, FIELD_GRID_XYZ = 0, 1
SMOOTHER_GAIN_MATRIX, SMOOTHER_MBF = 0, 1
STEPPER_EIGEN, STEPPER_SYMPY = 0, 1
# track-state table (acts_kf_track_states_device): parameter sets, float and
# int column indices
TS_PREDICTED, TS_FILTERED, TS_SMOOTHED, TS_UNBIASED = 0, 1, 2, 3
TS_CHI2, TS_PATH_LENGTH, TS_T_X, TS_T_PARAMS, TS_L_X_HIT, TS_RES_X_HIT = 0, 1, 2, 9, 15, 20
TS_NFLOAT = 154
TS_STATE_TYPE, TS_DIM_HIT, TS_HAS_PARAMS, TS_TRACK = 0, 1, 2, 3
TS_NINT = 4


def ts_params(p: int) -> int:
    """First float column of parameter set p (ACTS_KF_TS_PARAMS)."""
    return 26 + 32 * p
PRECISION_FP64, PRECISION_MIXED = 0, 1


class Surface(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("material", C.c_int32),
        ("flags", C.c_uint32),
        ("reserved", C.c_int32),
        ("rotation", C.c_double * 9),
        ("translation", C.c_double * 3),
        ("bounds", C.c_double * 2),
    ]


class MaterialSlab(C.Structure):
    _fields_ = [
        ("x0", C.c_float),
        ("l0", C.c_float),
        ("ar", C.c_float),
        ("z", C.c_float),
        ("molar_rho", C.c_float),
        ("thickness", C.c_float),
    ]


class Field(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("reserved", C.c_int32),
        ("constant", C.c_double * 3),
        ("n", C.c_int32 * 3),
        ("reserved2", C.c_int32),
        ("point_min", C.c_double * 3),
        ("point_max", C.c_double * 3),
        ("values", C.POINTER(C.c_double)),
    ]


class Options(C.Structure):
    _fields_ = [
        ("multiple_scattering", C.c_int32),
        ("energy_loss", C.c_int32),
        ("use_joseph", C.c_int32),
        ("precision", C.c_int32),
        ("smoother", C.c_int32),
        ("stepper", C.c_int32),
        ("outlier_chi2_cut", C.c_double),
        ("step_tolerance", C.c_double),
        ("step_size_cutoff", C.c_double),
        ("initial_step_size", C.c_double),
        ("max_step_size", C.c_double),
        ("max_runge_kutta_step_trials", C.c_uint32),
        ("max_steps", C.c_uint32),
        ("max_target_skipping", C.c_uint32),
        ("loop_protection", C.c_int32),
        ("path_limit", C.c_double),
        ("loop_fraction", C.c_double),
        ("surface_tolerance", C.c_double),
        ("near_limit", C.c_double),
        ("abs_pdg", C.c_int32),
        ("mass", C.c_float),
        ("abs_charge", C.c_float),
        ("reference_surface", C.c_int32),
        ("reference_strategy", C.c_int32),
        ("reference_position", C.c_double * 3),
        ("length_sorted", C.c_int32),
        ("sched_inner_steps", C.c_int32),
        ("sched_wait_64ths", C.c_int32),
        ("sched_quarantine_steps", C.c_int32),
        ("kernel_timing", C.c_int32),
        ("use_graph", C.c_int32),
        ("sched_transport_skip_steps", C.c_int32),
        ("navigator", C.c_int32),
        ("reverse_filtering", C.c_int32),
        ("reserved_reverse", C.c_int32),
        ("reverse_filtering_cov_scaling", C.c_double),
        ("reverse_filtering_momentum_threshold", C.c_double),
    ]


_P = C.c_void_p


class BatchIn(C.Structure):
    _fields_ = [
        ("n_tracks", C.c_int64),
        ("start_params", _P),
        ("start_cov", _P),
        ("start_plane", _P),
        ("seq_offsets", _P),
        ("seq_surfaces", _P),
        ("meas_offsets", _P),
        ("meas_surfaces", _P),
        ("meas_values", _P),
        ("meas_cov", _P),
        ("max_sequence_length", C.c_int32),
        ("reserved", C.c_int32),
        ("meas_dims", _P),
        ("meas_indices", _P),
    ]


BATCH_OUT_FIELDS = [
    "status", "n_states", "chi2", "ndf", "n_measurements", "n_holes",
    "n_outliers", "n_steps", "n_step_trials", "state_surface", "state_flags",
    "state_path", "state_chi2", "predicted", "predicted_cov", "filtered",
    "filtered_cov", "smoothed", "smoothed_cov", "jacobian", "reference_params",
    "reference_cov", "state_measurement", "reversed",
]


class BatchOut(C.Structure):
    _fields_ = [(name, _P) for name in BATCH_OUT_FIELDS]


class TrackStatesOut(C.Structure):
    _fields_ = [("n_slots", C.c_int64), ("columns", _P), ("ints", _P)]


class GenConfig(C.Structure):
    _fields_ = [
        ("n_tracks", C.c_int64),
        ("seed", C.c_uint64),
        ("pt_min", C.c_double),
        ("pt_max", C.c_double),
        ("pt_log_uniform", C.c_int32),
        ("n_threads", C.c_int32),
        ("eta_min", C.c_double),
        ("eta_max", C.c_double),
        ("phi_min", C.c_double),
        ("phi_max", C.c_double),
        ("vertex_sigma_z", C.c_double),
        ("sigma_loc0", C.c_double),
        ("sigma_loc1", C.c_double),
        ("start_sigma", C.c_double * 6),
        ("hole_probability", C.c_double),
        ("outlier_fraction", C.c_double),
        ("outlier_min", C.c_int32),
        ("outlier_max", C.c_int32),
        ("outlier_distance", C.c_double),
        ("min_measurements", C.c_int32),
        ("max_attempts", C.c_int32),
    ]


class Solenoid(C.Structure):
    _fields_ = [
        ("radius", C.c_double),
        ("length", C.c_double),
        ("n_coils", C.c_int32),
        ("reserved", C.c_int32),
        ("b_mag_center", C.c_double),
    ]


# symbol -> (restype, argtypes); every declaration in include/*.h
SIGNATURES = {
    "acts_kf_default_options": (None, [C.POINTER(Options)]),
    "acts_kf_create": (C.c_int, [C.POINTER(Surface), C.c_int32, C.POINTER(MaterialSlab),
                                 C.c_int32, C.POINTER(Field), C.POINTER(Options),
                                 C.c_int32, C.POINTER(C.c_void_p)]),
    "acts_kf_fit_batch": (C.c_int, [C.c_void_p, C.POINTER(BatchIn), C.POINTER(BatchOut),
                                    C.c_void_p]),
    "acts_kf_fit_batch_device": (C.c_int, [C.c_void_p, C.POINTER(BatchIn),
                                           C.POINTER(BatchOut), C.c_void_p]),
    "acts_kf_fit_batch_device_noalloc": (C.c_int, [C.c_void_p, C.POINTER(BatchIn),
                                                   C.POINTER(BatchOut), C.c_void_p]),
    "acts_kf_sync": (C.c_int, [C.c_void_p]),
    "acts_kf_kernel_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]),
    "acts_kf_track_states_device": (C.c_int, [C.c_void_p, C.POINTER(BatchIn),
                                              C.POINTER(BatchOut), C.c_void_p,
                                              C.POINTER(TrackStatesOut), C.c_void_p]),
    "acts_kf_destroy": (None, [C.c_void_p]),
    "acts_kf_status_string": (C.c_char_p, [C.c_int32]),
    "acts_kf_last_error": (C.c_char_p, []),
    "acts_kf_gen_run": (C.c_int, [C.POINTER(GenConfig), C.POINTER(Surface), C.c_int32,
                                  C.POINTER(MaterialSlab), C.c_int32, C.POINTER(Field),
                                  C.POINTER(Options), C.POINTER(C.c_void_p)]),
    "acts_kf_gen_sizes": (None, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int64)]),
    "acts_kf_gen_fetch": (None, [C.c_void_p] + [C.c_void_p] * 12),
    "acts_kf_gen_free": (None, [C.c_void_p]),
    "acts_kf_solenoid_map": (C.c_int, [C.POINTER(Solenoid), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.c_int32]),
}

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libacts_kf.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} not found: build it with `python -m acts_amd.build` "
            "(there is no CPU fallback)")
    # PyTorch for ROCm bundles its own HIP runtime and links it as
    # libamdhip64.so, a name the library's dependency (libamdhip64.so.7)
    # does not match: imported after the library, torch maps a second
    # runtime into the process, and whichever initialises second finds no
    # device.  Imported first, its runtime (soname libamdhip64.so.7) is the
    # one the library binds to.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def default_options() -> Options:
    o = Options()
    load_library().acts_kf_default_options(C.byref(o))
    return o


def last_error() -> str:
    return load_library().acts_kf_last_error().decode()


def status_string(code: int) -> str:
    return load_library().acts_kf_status_string(int(code)).decode()


def check(code: int, what: str) -> None:
    if code != OK:
        raise RuntimeError(f"{what} failed ({code}): {last_error()}")
