/*
 * acts_kf_gen.h -- synthetic event generator for the batched Kalman fitter.
 *
 * Produces the inputs of acts_kf_fit_batch for muons in a detector given as
 * an acts_kf_surface table: truth particles are propagated with an adaptive
 * Runge-Kutta-Nystroem integrator through the field, crossing every surface
 * they reach in order, losing the mean (Bethe) energy in each material slab
 * and scattering only through the covariance (no random scattering), as the
 * reference's MeasurementsCreator + MaterialInteractor truth propagation
 * does (Tests/CommonHelpers/.../MeasurementsCreator.hpp:51-174,
 * Core/include/Acts/Propagator/MaterialInteractor.hpp).  Measurements are the
 * truth local positions smeared with Gaussian resolutions; start parameters
 * are the truth parameters on a curvilinear plane at the vertex, smeared.
 *
 * Every track draws from its own counter-based random stream (seed, track
 * index), so the output does not depend on the number of threads.
 */
#ifndef ACTS_KF_GEN_H
#define ACTS_KF_GEN_H

#include <stdint.h>

#include "acts_kf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct acts_kf_gen_config {
  int64_t n_tracks;
  uint64_t seed;
  double pt_min, pt_max;   /* GeV */
  int32_t pt_log_uniform;  /* 0: uniform in pT, 1: uniform in log pT */
  int32_t n_threads;       /* <= 0: all cores */
  double eta_min, eta_max;
  double phi_min, phi_max;
  double vertex_sigma_z;   /* mm */
  double sigma_loc0;       /* mm: measurement resolution, first local coordinate */
  double sigma_loc1;       /* mm: second local coordinate (on discs: r*phi) */
  /* start-parameter smearing: loc0 [mm], loc1 [mm], phi, theta [rad],
   * relative q/p, time [mm] */
  double start_sigma[6];
  double hole_probability;    /* per crossed sensitive surface */
  double outlier_fraction;    /* fraction of tracks carrying outliers */
  int32_t outlier_min, outlier_max; /* outliers per such track */
  double outlier_distance;    /* in units of the measurement sigma */
  int32_t min_measurements;   /* tracks with fewer are regenerated */
  int32_t max_attempts;       /* regeneration attempts per track */
} acts_kf_gen_config;

typedef struct acts_kf_gen acts_kf_gen;

/* Generate; afterwards query the sizes with acts_kf_gen_sizes. */
int acts_kf_gen_run(const acts_kf_gen_config* cfg, const acts_kf_surface* surfaces,
                    int32_t n_surfaces, const acts_kf_material_slab* materials,
                    int32_t n_materials, const acts_kf_field* field,
                    const acts_kf_options* particle, acts_kf_gen** out);

/* total number of sequence entries and measurements */
void acts_kf_gen_sizes(const acts_kf_gen* g, int64_t* n_tracks, int64_t* n_seq,
                       int64_t* n_meas);

/* Copy the batch input arrays (sizes from acts_kf_gen_sizes) and the truth:
 * truth_start [n][6] (bound parameters on the start plane), truth_meas
 * [n_meas][6] (bound parameters at each measurement, before the surface's
 * material), meas_is_outlier [n_meas].  Any pointer may be NULL.           */
void acts_kf_gen_fetch(const acts_kf_gen* g, double* start_params, double* start_cov,
                       double* start_plane, int64_t* seq_offsets, int32_t* seq_surfaces,
                       int64_t* meas_offsets, int32_t* meas_surfaces, double* meas_values,
                       double* meas_cov, double* truth_start, double* truth_meas,
                       int32_t* meas_is_outlier);

void acts_kf_gen_free(acts_kf_gen* g);

/* ---- field maps -----------------------------------------------------------
 * SolenoidBField::Config (Core/include/Acts/MagneticField/SolenoidBField.hpp):
 * a finite solenoid of n_coils current loops, scaled so that the field at
 * the origin is b_mag_center.                                               */
typedef struct acts_kf_solenoid {
  double radius;        /* mm */
  double length;        /* mm */
  int32_t n_coils;
  int32_t reserved;
  double b_mag_center;  /* field at the centre, GeV/(e mm) */
} acts_kf_solenoid;

/* Sample SolenoidBField::getField (Core/src/MagneticField/SolenoidBField.cpp:
 * 24-162, including its use of k^2 as the elliptic modulus) on the points of
 * a regular grid, point i of axis a at point_min[a] + i * (point_max[a] -
 * point_min[a]) / (n[a] - 1), into values[((ix*n[1]) + iy)*n[2] + iz][3]
 * (the layout of acts_kf_field).  The fieldMapXYZ (BFieldMapUtils.cpp:
 * 109-186) step of the reference's map construction.                        */
int acts_kf_solenoid_map(const acts_kf_solenoid* solenoid, const int32_t n[3],
                         const double point_min[3], const double point_max[3],
                         double* values, int32_t n_threads);

#ifdef __cplusplus
}
#endif

#endif /* ACTS_KF_GEN_H */
