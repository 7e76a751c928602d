/*
 * acts_kf.h -- C ABI of the MI355X batched Kalman track fitter.
 *
 * This is the drop-in boundary for the reference call
 *
 *   Acts::KalmanFitter<Propagator<EigenStepper<>, DirectNavigator>, traj_t>::fit(
 *       begin, end, startParameters, kfOptions, surfaceSequence, trackContainer)
 *   (Core/include/Acts/TrackFitting/KalmanFitter.hpp:721-731)
 *
 * batched over many tracks.  Everything is plain C: POD structs, pointers and
 * sizes.  One call fits a whole batch of independent tracks; per-track results
 * and error codes come back in structure-of-arrays buffers.
 *
 * Conventions (identical to the reference):
 *   - units: mm, GeV, ns*c (time in mm), magnetic field in GeV/(e*mm)
 *     (1 T = 0.000299792458, Core/include/Acts/Definitions/Units.hpp:143)
 *   - bound parameter order: loc0, loc1, phi, theta, q/p, time
 *     (Core/include/Acts/Definitions/TrackParametrization.hpp:42-53)
 *   - free parameter order: x, y, z, t, dx, dy, dz, q/p (:67-79)
 *   - every 6x6 matrix (covariance, jacobian) is stored row-major, 36 doubles.
 *
 * Ownership: the caller owns every buffer it passes.  Host buffers handed to
 * acts_kf_fit_batch must stay valid until acts_kf_sync returns.  The context
 * owns its device memory (geometry, field, scratch).  One context serialises
 * its calls on the stream it is given; several contexts may share a device.
 */
#ifndef ACTS_KF_H
#define ACTS_KF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 2: acts_kf_batch_in gained meas_dims and meas_indices (measurement
 * subspaces); callers that value-initialise the struct are unaffected.
 * 3: acts_kf_options gained the reverse-filtering options and
 * acts_kf_batch_out the per-track `reversed` flag (both appended).         */
#define ACTS_KF_API_VERSION 4

/* ---- API return codes (the call itself) -------------------------------- */
#define ACTS_KF_OK 0
#define ACTS_KF_ERR_INVALID_ARGUMENT (-1)
#define ACTS_KF_ERR_DEVICE (-2)         /* a HIP runtime call failed */
#define ACTS_KF_ERR_OUT_OF_MEMORY (-3)
#define ACTS_KF_ERR_UNSUPPORTED (-4)

/* ---- per-track fit status: the reference's error enums ----------------- *
 * code = category | enum value, enum values as in the reference headers.   */
#define ACTS_KF_STATUS_OK 0
#define ACTS_KF_CAT_KALMAN 0x100     /* KalmanFitterError.hpp:18-31 */
#define ACTS_KF_CAT_PROPAGATOR 0x200 /* PropagatorError.hpp:18-26 */
#define ACTS_KF_CAT_STEPPER 0x300    /* EigenStepperError.hpp:18-26 */
#define ACTS_KF_CAT_SURFACE 0x400    /* SurfaceError.hpp:18-22 */
#define ACTS_KF_CAT_FIELD 0x500      /* MagneticFieldError.hpp:18-22 */
#define ACTS_KF_CAT_EXTRAPOLATION 0x600 /* TrackExtrapolationError, TrackHelpers.hpp:48-53 */
#define ACTS_KF_CAT_LIBRARY 0x700    /* this library's own conditions */

#define ACTS_KF_STATUS_UPDATE_FAILED (ACTS_KF_CAT_KALMAN | 1)
#define ACTS_KF_STATUS_SMOOTH_FAILED (ACTS_KF_CAT_KALMAN | 2)
#define ACTS_KF_STATUS_NO_MEASUREMENT_FOUND (ACTS_KF_CAT_KALMAN | 4)
#define ACTS_KF_STATUS_REVERSE_PROPAGATION_FAILED (ACTS_KF_CAT_KALMAN | 5)
#define ACTS_KF_STATUS_INCONSISTENT_TRACK_STATES (ACTS_KF_CAT_KALMAN | 6)
#define ACTS_KF_STATUS_PROPAGATION_FAILURE (ACTS_KF_CAT_PROPAGATOR | 1)
#define ACTS_KF_STATUS_STEP_COUNT_LIMIT (ACTS_KF_CAT_PROPAGATOR | 2)
#define ACTS_KF_STATUS_NEXT_TARGET_LIMIT (ACTS_KF_CAT_PROPAGATOR | 3)
#define ACTS_KF_STATUS_STEP_SIZE_STALLED (ACTS_KF_CAT_STEPPER | 1)
#define ACTS_KF_STATUS_STEP_INVALID (ACTS_KF_CAT_STEPPER | 2)
#define ACTS_KF_STATUS_STEP_ADJUSTMENT_FAILED (ACTS_KF_CAT_STEPPER | 3)
#define ACTS_KF_STATUS_NOT_ON_SURFACE (ACTS_KF_CAT_SURFACE | 1)
#define ACTS_KF_STATUS_FIELD_OUT_OF_BOUNDS (ACTS_KF_CAT_FIELD | 1)
#define ACTS_KF_STATUS_NO_COMPATIBLE_STATE (ACTS_KF_CAT_EXTRAPOLATION | 1)
#define ACTS_KF_STATUS_REFERENCE_UNREACHABLE (ACTS_KF_CAT_EXTRAPOLATION | 2)
/* the layer navigator found more states than the track has slots
 * (seq_offsets[i+1] - seq_offsets[i]); the reference's track container
 * grows instead, so it has no such error                                   */
#define ACTS_KF_STATUS_STATE_CAPACITY (ACTS_KF_CAT_LIBRARY | 1)

/* ---- track-state type flags (TrackStateType.hpp:30-49, bit positions) -- */
#define ACTS_KF_STATE_HAS_MEASUREMENT (1u << 0)
#define ACTS_KF_STATE_HAS_PARAMETERS (1u << 1)
#define ACTS_KF_STATE_IS_OUTLIER (1u << 2)
#define ACTS_KF_STATE_IS_HOLE (1u << 3)
#define ACTS_KF_STATE_HAS_MATERIAL (1u << 4)

/* ---- geometry ----------------------------------------------------------- */
#define ACTS_KF_SURFACE_PLANE 0    /* PlaneSurface.cpp */
#define ACTS_KF_SURFACE_CYLINDER 1 /* CylinderSurface.cpp */
#define ACTS_KF_SURFACE_DISC 2     /* DiscSurface.cpp */
/* LineSurface.cpp; only as the reference surface (PerigeeSurface), not in
 * the geometry table */
#define ACTS_KF_SURFACE_LINE 3

/* Reference surface of the fitted track parameters
 * (KalmanFitterOptions::referenceSurface, KalmanFitter.hpp:159-163,959-971):
 * none, or a PerigeeSurface (a line along the global z axis) through
 * acts_kf_options.reference_position, as the Examples set it up
 * (TrackFittingAlgorithm.cpp:89-90).                                        */
#define ACTS_KF_REFERENCE_NONE 0
#define ACTS_KF_REFERENCE_PERIGEE 1
/* TrackExtrapolationStrategy (TrackHelpers.hpp:36-44) */
#define ACTS_KF_EXTRAPOLATE_FIRST 0
#define ACTS_KF_EXTRAPOLATE_LAST 1
#define ACTS_KF_EXTRAPOLATE_FIRST_OR_LAST 2

#define ACTS_KF_SURFACE_SENSITIVE 1u /* Surface::isSensitive() */

/* Navigator of the fit (acts_kf_options.navigator).
 * DIRECT: DirectNavigator through each track's given surface sequence
 *   (KalmanFitter::fit with sSequence, KalmanFitter.hpp:725,778).
 * LAYERS: the surfaces are found from the geometry (KalmanFitter::fit
 *   without a sequence, KalmanFitter.hpp:693-699, with the Navigator as the
 *   reference's fitter tests configure it: resolveSensitive and
 *   resolveMaterial on, resolvePassive off, FitterTestsCommon.hpp:96-99).
 *   Every surface of the geometry table that is sensitive or carries
 *   material is a layer of its own (Navigator Gen1, Navigator.cpp:251-316,
 *   479-540,826-849).  At the start and after each reached layer, the
 *   layers are resolved along the straight line of the current direction,
 *   without the layer the track is on: the first intersection of each,
 *   inside its bounds, more than the on-surface tolerance ahead, sorted by
 *   path length (TrackingVolume::compatibleLayers, Layer::
 *   surfaceOnApproach); the propagation targets them in that order,
 *   skipping those it can no longer reach inside their bounds, and once
 *   they are used up resolves again from where the track is
 *   (renavigation); when that finds none the navigation ends.  (Gen1
 *   resolves a volume's layers once, and again after a volume switch; here
 *   each reached layer counts as one.)  seq_offsets still give every track its state
 *   slots (its capacity); the entries of seq_surfaces are not read.  The
 *   reverse filter navigates the same way backwards from the last
 *   measurement surface.  The transport-skip option is ignored.            */
#define ACTS_KF_NAVIGATOR_DIRECT 0
#define ACTS_KF_NAVIGATOR_LAYERS 1

/* One surface.  The placement is the surface's local-to-global transform:
 * rotation columns are the local x, y, z axes expressed in the global frame.
 * The direct navigator does not use the bounds (it intersects with
 * BoundaryTolerance::Infinite, DirectNavigator.hpp:305-308); the layer
 * navigator does (BoundaryTolerance::None), as do the geometry helpers and
 * the event generator.  Plane bounds: half lengths of a RectangleBounds in
 * local x and y, or 0 and 0 for an unbounded plane.                          */
typedef struct acts_kf_surface {
  int32_t type;          /* ACTS_KF_SURFACE_* */
  int32_t material;      /* index into the material table, or -1 */
  uint32_t flags;        /* ACTS_KF_SURFACE_* flag bits */
  int32_t reserved;
  double rotation[9];    /* row-major 3x3 */
  double translation[3];
  double bounds[2];      /* cylinder: R, half length z; disc: rmin, rmax */
} acts_kf_surface;

/* Homogeneous surface material: Material::fromMolarDensity(x0, l0, ar, z,
 * molar_rho) with a slab thickness (Material.cpp, MaterialSlab.hpp:40-160).
 * The library derives the electron density and mean excitation energy in
 * float arithmetic exactly as the reference does.                          */
typedef struct acts_kf_material_slab {
  float x0;         /* radiation length [mm] */
  float l0;         /* nuclear interaction length [mm] */
  float ar;         /* relative atomic mass */
  float z;          /* nuclear charge */
  float molar_rho;  /* molar density [mol/mm^3] */
  float thickness;  /* [mm] */
} acts_kf_material_slab;

/* ---- magnetic field ------------------------------------------------------ */
#define ACTS_KF_FIELD_CONSTANT 0 /* ConstantBField.hpp */
#define ACTS_KF_FIELD_GRID_XYZ 1 /* fieldMapXYZ + InterpolatedBFieldMap */

typedef struct acts_kf_field {
  int32_t kind;
  int32_t reserved;
  double constant[3];  /* kind == CONSTANT */
  /* kind == GRID_XYZ: field values on the points of a regular grid,
   * n[a] points on axis a from point_min[a] to point_max[a] (inclusive).
   * values[((ix*n[1]) + iy)*n[2] + iz][c], c = Bx, By, Bz.                  */
  int32_t n[3];
  int32_t reserved2;
  double point_min[3];
  double point_max[3];
  const double* values; /* host pointer, read once by acts_kf_create */
} acts_kf_field;

/* ---- options: KalmanFitterOptions + StepperPlainOptions +
 *      PropagatorPlainOptions + DirectNavigator::Options + particle -------- */
#define ACTS_KF_PRECISION_FP64 0  /* plain FP64 fit (the drop-in) */
#define ACTS_KF_PRECISION_MIXED 1 /* FP32 stepping/field, FP64 Kalman algebra */
/* KalmanFitterExtensions::smoother: GainMatrixSmoother (RTS,
 * GainMatrixSmoother.cpp:18-97) or MbfSmoother (modified Bryson-Frazier,
 * MbfSmoother.cpp, detail/MbfSmootherImpl.hpp), which the Examples use
 * (KalmanFitterFunction.cpp:39-44,100-102)                                 */
/* the stepper of the propagator: EigenStepper (EigenStepper.ipp:169-388,
 * the drop-in) or SympyStepper (Core/src/Propagator/SympyStepper.cpp:
 * 141-305), the Examples' stepper (KalmanFitterFunction.cpp:39-44), whose
 * vacuum step carries the bound-to-free jacobian through an ATLAS-form RK4
 * (Core/src/Propagator/generate_sympy_stepper.py:157-444); with
 * ACTS_KF_PRECISION_MIXED its stages and jacobian run in FP32.              */
#define ACTS_KF_STEPPER_EIGEN 0
#define ACTS_KF_STEPPER_SYMPY 1
#define ACTS_KF_SMOOTHER_GAIN_MATRIX 0
#define ACTS_KF_SMOOTHER_MBF 1

typedef struct acts_kf_options {
  int32_t multiple_scattering; /* KalmanFitter.hpp:166, default 1 */
  int32_t energy_loss;         /* KalmanFitter.hpp:169, default 1 */
  int32_t use_joseph;          /* GainMatrixUpdater, default 0 */
  int32_t precision;           /* ACTS_KF_PRECISION_* */
  int32_t smoother;            /* ACTS_KF_SMOOTHER_*, default GAIN_MATRIX */
  int32_t stepper;             /* ACTS_KF_STEPPER_*, default EIGEN */
  double outlier_chi2_cut;     /* predicted chi2 cut; <= 0 disables the finder */
  /* StepperPlainOptions (StepperOptions.hpp:36-49) */
  double step_tolerance;       /* 1e-4 */
  double step_size_cutoff;     /* 0 */
  double initial_step_size;    /* 10 m */
  double max_step_size;        /* DBL_MAX */
  uint32_t max_runge_kutta_step_trials; /* 10000 */
  /* PropagatorPlainOptions (PropagatorOptions.hpp:26-52) */
  uint32_t max_steps;          /* 1000 */
  uint32_t max_target_skipping;/* 100 */
  int32_t loop_protection;     /* 1 */
  double path_limit;           /* DBL_MAX */
  double loop_fraction;        /* 0.5 */
  double surface_tolerance;    /* 1e-4 */
  /* DirectNavigator::Options (DirectNavigator.hpp:51-61) */
  double near_limit;           /* -100 um */
  /* ParticleHypothesis (ParticleHypothesis.hpp) */
  int32_t abs_pdg;             /* 13 = muon */
  float mass;                  /* GeV */
  float abs_charge;            /* e */
  /* extrapolation of the smoothed track to a reference surface after the
   * fit (extrapolateTrackToReferenceSurface, TrackHelpers.hpp:339-383):
   * ACTS_KF_REFERENCE_*, ACTS_KF_EXTRAPOLATE_* (default first-or-last, as
   * KalmanFitterOptions; the Examples use first), and the perigee point.
   * The propagation uses the reference's default propagator and stepper
   * options, not the ones above (KalmanFitter.hpp:960-961).                 */
  int32_t reference_surface;
  int32_t reference_strategy;
  double reference_position[3];
  /* scheduling (no reference counterpart; results do not depend on it):
   * length_sorted     order in which the kernels take the tracks: 1 =
   *                   decreasing sequence length (default), 0 = input
   *                   order, 2 = by start direction and charge (bins of
   *                   eta, phi and the sign of q/p, so that neighbouring
   *                   lanes follow neighbouring paths), 3 = by sequence
   *                   length, then by start momentum (bins of log p, so
   *                   that neighbouring lanes bend alike); tracks with a
   *                   measurement near the +-pi seam of a cylinder first;
   * sched_inner_steps most RK steps the forward kernel takes before the
   *                   lanes waiting at a surface are updated (0 = 64);
   * sched_wait_64ths  the surface update runs early once this many 64ths of
   *                   the busy lanes of a wavefront wait (0 = 64: all);
   * sched_quarantine_steps  a wavefront takes no new tracks while one of its
   *                   tracks has run more than this many steps (0 = off).   */
  int32_t length_sorted;
  int32_t sched_inner_steps;
  int32_t sched_wait_64ths;
  int32_t sched_quarantine_steps;
  /* 1 = record HIP events around the kernels of each fit, for
   * acts_kf_kernel_times (no reference counterpart)                         */
  int32_t kernel_timing;
  /* 1 = capture the launch sequence of a fit in a hipGraph and replay it for
   * later fits with the same buffers and sizes (no reference counterpart;
   * kernel_timing is ignored then)                                           */
  int32_t use_graph;
  /* a track that has taken this many steps since its last surface saves its
   * state and steps on without the jacobian transport; should it reach a
   * surface after all, it goes back to the saved state and repeats those
   * steps with the transport, so the results do not depend on it (a track
   * that never reaches another surface, e.g. a looper that ends at
   * max_steps, uses no transport).  0 = off (default): a track that does
   * reach a surface repeats the skipped steps, which costs more than the
   * loopers save in the field map.  No reference counterpart.               */
  int32_t sched_transport_skip_steps;
  /* ACTS_KF_NAVIGATOR_* (default DIRECT)                                    */
  int32_t navigator;
  /* Reverse filtering instead of smoothing (KalmanFitter.hpp:884-945,
   * KalmanFitterOptions::reverseFiltering and
   * reverseFilteringCovarianceScaling, :171-181): after the forward filter
   * the track is filtered again in the opposite direction, from the last
   * measurement state's filtered parameters with the covariance times
   * reverse_filtering_cov_scaling (default 100), through the same surface
   * sequence, with the reference surface as its target.  The first
   * measurement state's smoothed parameters are the reverse filter's
   * filtered parameters on that surface, the last measurement state's are
   * its own filtered parameters; no other state has smoothed parameters
   * (their smoothed slots hold NaN).  A track is filtered in reverse when
   * reverse_filtering is 1, or when |1 / filtered q/p| of its last
   * measurement state is at most reverse_filtering_momentum_threshold
   * (the Examples' SimpleReverseFilteringLogic, KalmanFitterFunction.cpp:
   * 53-57; default 0: never).                                              */
  int32_t reverse_filtering;
  int32_t reserved_reverse;
  double reverse_filtering_cov_scaling;
  double reverse_filtering_momentum_threshold;
} acts_kf_options;

/* ---- one batch of tracks ------------------------------------------------- */
typedef struct acts_kf_batch_in {
  int64_t n_tracks;
  /* start parameters: bound vector and covariance on the track's start plane
   * (BoundTrackParameters; the start surface is a PlaneSurface given by its
   * rotation[9] + translation[3] per track, without material)               */
  const double* start_params;   /* [n][6] */
  const double* start_cov;      /* [n][36] */
  const double* start_plane;    /* [n][12]: rotation row-major, then translation */
  /* surface sequence for the direct navigator (KalmanFitter.hpp:725,778);
   * with the layer navigator only the offsets are used (each track's state
   * slots), but seq_surfaces must still point to seq_offsets[n] entries   */
  const int64_t* seq_offsets;   /* [n+1] */
  const int32_t* seq_surfaces;  /* [seq_offsets[n]] indices into the geometry */
  /* measurements: 2D (loc0, loc1) source links keyed by their surface */
  const int64_t* meas_offsets;  /* [n+1] */
  const int32_t* meas_surfaces; /* [meas_offsets[n]] */
  const double* meas_values;    /* [m][2] */
  const double* meas_cov;       /* [m][4] row-major 2x2 */
  /* upper bound on any track's sequence length (sizes the per-state scratch);
   * 0 = let the library compute it (one extra device->host copy)            */
  int32_t max_sequence_length;
  int32_t reserved;
  /* Measurement subspaces, optional (NULL: every measurement is 2D on
   * (loc0, loc1)).  meas_dims[m] = N in {1, 2}; meas_indices[m][0..N-1] =
   * the bound parameters measured (0..5, distinct); the values and the N x N
   * row-major covariance fill the first N and N*N entries of meas_values[m]
   * and meas_cov[m] (a strip measurement of loc1: dims 1, indices {1, -},
   * meas_values[m][0] = loc1, meas_cov[m][0] = its variance).  The
   * reference's GainMatrixUpdater dispatches on N = 1..6 with the
   * projector of the subspace indices (Core/src/TrackFitting/
   * GainMatrixUpdater.cpp:19-29, detail/GainMatrixUpdaterImpl.hpp:21-91);
   * N above 2 is not supported here.  acts_kf_fit_batch rejects other
   * values with ACTS_KF_ERR_INVALID_ARGUMENT; on the device path a track
   * with such a measurement fails with ACTS_KF_STATUS_PROPAGATION_FAILURE. */
  const uint8_t* meas_dims;     /* [m] or NULL */
  const uint8_t* meas_indices;  /* [m][2] or NULL */
} acts_kf_batch_in;

/* Output buffers.  Per-state arrays use the sequence layout: the states of
 * track i occupy slots seq_offsets[i] .. seq_offsets[i] + n_states[i] - 1 in
 * propagation order (first state first), up to and including the last
 * measurement state (the track tip, KalmanFitter.hpp:831).  The remaining
 * slots of the track (all of them when its status is not 0) are written as
 * empty slots: state_surface and state_measurement -1, state_flags 0 and
 * every number 0.  Any per-state pointer may be NULL to skip that output.  */
typedef struct acts_kf_batch_out {
  int32_t* status;          /* [n] ACTS_KF_STATUS_* */
  int32_t* n_states;        /* [n] */
  float* chi2;              /* [n] sum over measurement states (float, as the track column) */
  int32_t* ndf;             /* [n] */
  int32_t* n_measurements;  /* [n] */
  int32_t* n_holes;         /* [n] */
  int32_t* n_outliers;      /* [n] */
  int32_t* n_steps;         /* [n] accepted RK steps (StepperStatistics) */
  int32_t* n_step_trials;   /* [n] attempted RK steps */
  int32_t* state_surface;   /* [slots] geometry index of the state's surface */
  uint32_t* state_flags;    /* [slots] ACTS_KF_STATE_* */
  double* state_path;       /* [slots] accumulated path length */
  float* state_chi2;        /* [slots] filtered chi2 (float, VectorMultiTrajectory.hpp:193) */
  double* predicted;        /* [slots][6] */
  double* predicted_cov;    /* [slots][36] */
  double* filtered;         /* [slots][6] */
  double* filtered_cov;     /* [slots][36] */
  double* smoothed;         /* [slots][6] */
  double* smoothed_cov;     /* [slots][36] */
  double* jacobian;         /* [slots][36] transport jacobian from the previous state */
  /* per track, with a reference surface: the track parameters and
   * covariance on it (TrackProxy::parameters()/covariance())              */
  double* reference_params; /* [n][6] */
  double* reference_cov;    /* [n][36] */
  /* [slots] the state's measurement as an index into the batch's
   * measurement arrays (meas_surfaces/meas_values/meas_cov), i.e. its
   * uncalibrated source link (IndexSourceLink::index); -1 for hole and
   * material-only states                                                   */
  int32_t* state_measurement;
  /* [n] 1 if the track was smoothed by reverse filtering (the track's
   * "reversed" column, KalmanFitter.hpp:978-981), else 0                   */
  int32_t* reversed;
} acts_kf_batch_out;

typedef struct acts_kf_ctx acts_kf_ctx;

/* Fill *opt with the reference defaults (see the field comments): the
 * default-constructed KalmanFitterOptions / PropagatorPlainOptions /
 * StepperPlainOptions (KalmanFitter.hpp:118-175, PropagatorOptions.hpp:26-52,
 * StepperOptions.hpp:36-49) and ParticleHypothesis::muon().                */
void acts_kf_default_options(acts_kf_options* opt);

/* Create a fitting context on HIP device `device`: uploads the geometry,
 * material table and field once.  Replaces the construction of
 * KalmanFitter(Propagator(EigenStepper(bField), DirectNavigator())).       */
int acts_kf_create(const acts_kf_surface* surfaces, int32_t n_surfaces,
                   const acts_kf_material_slab* materials, int32_t n_materials,
                   const acts_kf_field* field, const acts_kf_options* options,
                   int32_t device, acts_kf_ctx** out_ctx);

/* Fit a batch whose buffers (in and out) are host memory.  Asynchronous on
 * `stream` (a hipStream_t, NULL = the context's own stream); results are in
 * the output buffers after acts_kf_sync.  seq_offsets and meas_offsets must
 * start at 0 and never decrease (ACTS_KF_ERR_INVALID_ARGUMENT otherwise).  Replaces one call of
 * KalmanFitter::fit (KalmanFitter.hpp:721-731, the DirectNavigator overload
 * with the surface sequence) per track of the batch, as issued by
 * KalmanFitterFunctionImpl (Examples/Algorithms/TrackFitting/src/
 * KalmanFitterFunction.cpp:150-157).                                        */
int acts_kf_fit_batch(acts_kf_ctx* ctx, const acts_kf_batch_in* in,
                      const acts_kf_batch_out* out, void* stream);

/* Fit a batch whose buffers are all device memory on the context's device
 * (inputs already resident in HBM).  Asynchronous on `stream`.  Same
 * reference call as acts_kf_fit_batch (KalmanFitter.hpp:721-731).
 * Precondition (not checked: the offsets are device memory): seq_offsets
 * and meas_offsets start at 0, never decrease, and their last entries are
 * the lengths of seq_surfaces and of the measurement arrays, which the
 * output slot arrays match.                                                 */
int acts_kf_fit_batch_device(acts_kf_ctx* ctx, const acts_kf_batch_in* in,
                             const acts_kf_batch_out* out, void* stream);

/* acts_kf_fit_batch_device without any device allocation: no counterpart in
 * the reference.  For callers that cannot allocate device memory during a
 * fit (code under hipGraph stream capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same inputs, outputs, preconditions and results
 * (bit for bit) as acts_kf_fit_batch_device; the fit's workspace (per-state
 * records, per-track results, queue, sort buffers, reference-surface
 * states) comes from a 256 MiB block of device memory that the library
 * holds on every device it runs on until the process ends.  A batch needs
 * 748 bytes per track and state slot (n_tracks x max_sequence_length) plus
 * about 16 (+ 432 with a reference surface) bytes per track; a larger batch
 * is rejected with ACTS_KF_ERR_INVALID_ARGUMENT before anything is queued.
 * Every call on a device shares that block: calls that overlap on one
 * device, from any context or thread, race (serialise them, e.g. on one
 * stream).  The fit is queued eagerly (options.use_graph is not used) and
 * without transport skipping (options.sched_transport_skip_steps, which
 * never changes results).  With max_sequence_length 0 the call waits for
 * the stream once to read the longest sequence back.                      */
int acts_kf_fit_batch_device_noalloc(acts_kf_ctx* ctx, const acts_kf_batch_in* in,
                                     const acts_kf_batch_out* out, void* stream);

/* Wait for the work of the context's last call (fit_batch, fit_batch_device
 * or track_states_device), on whichever stream it was queued; work of other
 * contexts on the device is not waited for.  The reference's fit() returns
 * with its result (KalmanFitter.hpp:721-731); a batched, asynchronous fit
 * makes that completion point explicit.                                    */
int acts_kf_sync(acts_kf_ctx* ctx);

/* Release the context and its device memory (the end of life of the
 * KalmanFitter / Propagator / field objects of the reference).             */
void acts_kf_destroy(acts_kf_ctx* ctx);

/* Durations in milliseconds of the forward-filter kernel and of the
 * smoother kernel of the context's last fit (needs options.kernel_timing;
 * waits for that fit).  No reference counterpart: a profiling aid.          */
int acts_kf_kernel_times(acts_kf_ctx* ctx, double* forward_ms, double* smoother_ms);

/* ---- track-state table ---------------------------------------------------
 * The per-state columns RootTrackStatesWriter derives from a fitted track
 * (Examples/Io/Root/src/RootTrackStatesWriter.cpp:364-735): for each of the
 * predicted, filtered, smoothed and unbiased parameters (the last from
 * calculateUnbiasedParametersCovariance, TrackHelpersUnbiased.cpp:21-54)
 * the values, their errors, residuals and pulls against the truth, the
 * global position and momentum, eta and pT; the hit's local and global
 * position and its residual, error and pull against the prediction; the
 * truth at the hit.  Values are computed in double and stored as float
 * through clampValue<float> (Helpers.hpp:170-188), NaN where the writer
 * pushes NaN.  One row per state slot (the sequence layout of
 * acts_kf_batch_out), stored column-major: column c of slot i is
 * columns[c * n_slots + i].  Rows of failed tracks and unused slots get
 * state type -1.  Where the writer pushes nothing for a state (residuals
 * and pulls of states without a measurement; hit residuals of states
 * without a measurement) the column holds NaN and the host-side writer
 * skips it.                                                                 */
#define ACTS_KF_TS_PREDICTED 0
#define ACTS_KF_TS_FILTERED 1
#define ACTS_KF_TS_SMOOTHED 2
#define ACTS_KF_TS_UNBIASED 3
/* float columns */
#define ACTS_KF_TS_CHI2 0        /* chi2 */
#define ACTS_KF_TS_PATH_LENGTH 1 /* pathLength */
#define ACTS_KF_TS_T_X 2         /* t_x t_y t_z t_r t_dx t_dy t_dz (2..8) */
#define ACTS_KF_TS_T_PARAMS 9    /* t_eLOC0 .. t_eT (9..14) */
#define ACTS_KF_TS_L_X_HIT 15    /* l_x_hit l_y_hit g_x_hit g_y_hit g_z_hit */
#define ACTS_KF_TS_RES_X_HIT 20  /* res_x_hit res_y_hit err_x_hit err_y_hit
                                    pull_x_hit pull_y_hit (20..25) */
/* per parameter set p (ACTS_KF_TS_PREDICTED..UNBIASED), 32 columns from
 * ACTS_KF_TS_PARAMS(p): eLOC0..eT (+0..5), res_ (+6..11), err_ (+12..17),
 * pull_ (+18..23), g_x g_y g_z (+24..26), px py pz (+27..29), eta (+30),
 * pT (+31)                                                                  */
#define ACTS_KF_TS_PARAMS(p) (26 + 32 * (p))
#define ACTS_KF_TS_NFLOAT 154
/* int columns */
#define ACTS_KF_TS_STATE_TYPE 0 /* stateType: 0 measurement, 1 outlier, 2 hole,
                                   3 material, 4 unknown; -1 no state */
#define ACTS_KF_TS_DIM_HIT 1    /* dim_hit (calibrated size, 0 without) */
#define ACTS_KF_TS_HAS_PARAMS 2 /* bit p set: parameter set p exists */
#define ACTS_KF_TS_TRACK 3      /* index of the state's track in the batch */
#define ACTS_KF_TS_NINT 4

typedef struct acts_kf_track_states_out {
  int64_t n_slots;  /* column stride, >= seq_offsets[n_tracks] */
  float* columns;   /* [ACTS_KF_TS_NFLOAT][n_slots] */
  int32_t* ints;    /* [ACTS_KF_TS_NINT][n_slots] */
} acts_kf_track_states_out;

/* Fill the track-state table of a fitted batch.  All buffers are device
 * memory on the context's device: `in` is the batch that was fitted,
 * `fitted` its outputs, which must hold status, n_states, state_surface,
 * state_flags, state_path, state_chi2, state_measurement and the
 * predicted, filtered and smoothed parameters and covariances.
 * `truth_meas` ([m][6] bound parameters of the truth particle at each
 * measurement, the writer's averaged sim hits) may be NULL: the truth
 * columns and the residuals and pulls are then NaN.  Asynchronous on
 * `stream`.  Replaces the per-state loop of RootTrackStatesWriter::writeT
 * (RootTrackStatesWriter.cpp:310-740); the host-side writer orders the
 * rows as the reference's trackStatesReversed() vectors.                    */
int acts_kf_track_states_device(acts_kf_ctx* ctx, const acts_kf_batch_in* in,
                                const acts_kf_batch_out* fitted, const double* truth_meas,
                                const acts_kf_track_states_out* out, void* stream);

/* Message for a per-track status code: the reference's error messages
 * (std::error_category::message of KalmanFitterError.cpp:21,
 * PropagatorError.cpp, EigenStepperError.cpp, SurfaceError.cpp,
 * MagneticFieldError.cpp).                                                  */
const char* acts_kf_status_string(int32_t status);

/* Message for the last API error of this thread (the reference reports
 * configuration errors by exceptions, VoidFitterComponents.hpp:22-27; a C
 * ABI returns a code and keeps the message here).                           */
const char* acts_kf_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* ACTS_KF_H */
