// oracle_capi.cpp -- C entry points of the CPU oracle.
// TEST INFRASTRUCTURE ONLY (see kf_oracle.hpp): loaded by tests/, by
// __graft_entry__.smoke() as the checker and by bench.py's cpu_baseline leg.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "kf_oracle.hpp"

using namespace acts_oracle;

namespace {

void buildGeometry(Geometry& geo, const acts_kf_surface* surfaces, int n,
                   const acts_kf_material_slab* mats, int nm,
                   const acts_kf_field* field, const acts_kf_options* opt) {
  geo.opt = *opt;
  // the extrapolation to the reference surface runs with default propagator
  // and stepper options (KalmanFitter.hpp:960-961; PropagatorOptions.hpp,
  // StepperOptions.hpp, DirectNavigator.hpp:51-61)
  geo.extrapolationOpt = *opt;
  geo.extrapolationOpt.step_tolerance = 1e-4;
  geo.extrapolationOpt.step_size_cutoff = 0.0;
  geo.extrapolationOpt.initial_step_size = 10000.0;
  geo.extrapolationOpt.max_step_size = std::numeric_limits<double>::max();
  geo.extrapolationOpt.max_runge_kutta_step_trials = 10000;
  geo.extrapolationOpt.max_steps = 1000;
  geo.extrapolationOpt.max_target_skipping = 100;
  geo.extrapolationOpt.loop_protection = 1;
  geo.extrapolationOpt.path_limit = std::numeric_limits<double>::max();
  geo.extrapolationOpt.loop_fraction = 0.5;
  geo.extrapolationOpt.surface_tolerance = 1e-4;
  geo.extrapolationOpt.near_limit = -0.1;
  geo.reference = Surface();
  geo.reference.type = ACTS_KF_SURFACE_LINE;
  {
    const double rot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    geo.reference.pl.set(rot, opt->reference_position);
  }
  geo.particle.absPdg = opt->abs_pdg;
  geo.particle.mass = opt->mass;
  geo.particle.absQ = opt->abs_charge;
  geo.surfaces.resize(n);
  for (int i = 0; i < n; ++i) {
    Surface& s = geo.surfaces[i];
    s.type = surfaces[i].type;
    s.pl.set(surfaces[i].rotation, surfaces[i].translation);
    s.b0 = surfaces[i].bounds[0];
    s.b1 = surfaces[i].bounds[1];
    s.sensitive = (surfaces[i].flags & ACTS_KF_SURFACE_SENSITIVE) != 0;
    const int mi = surfaces[i].material;
    s.hasMaterial = mi >= 0 && mi < nm;
    if (s.hasMaterial) {
      const acts_kf_material_slab& m = mats[mi];
      s.slab = SlabF::make(materialFromMolarDensity(m.x0, m.l0, m.ar, m.z, m.molar_rho),
                           m.thickness);
    }
  }
  geo.field.kind = field->kind;
  for (int i = 0; i < 3; ++i) geo.field.B[i] = field->constant[i];
  if (field->kind == ACTS_KF_FIELD_GRID_XYZ)
    geo.field.grid.set(field->n, field->point_min, field->point_max, field->values);
}

bool fieldSupported(const acts_kf_field* f) {
  if (f->kind == ACTS_KF_FIELD_CONSTANT) return true;
  if (f->kind != ACTS_KF_FIELD_GRID_XYZ || !f->values) return false;
  for (int a = 0; a < 3; ++a)
    if (f->n[a] < 2 || !(f->point_max[a] > f->point_min[a])) return false;
  return true;
}

void loadTrack(const acts_kf_batch_in* in, int64_t t, TrackIn& tr) {
  for (int i = 0; i < 6; ++i) tr.startParams.a[i][0] = in->start_params[6 * t + i];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) tr.startCov.a[i][j] = in->start_cov[36 * t + 6 * i + j];
  const double* p = in->start_plane + 12 * t;
  tr.startSurface = Surface();
  tr.startSurface.type = ACTS_KF_SURFACE_PLANE;
  tr.startSurface.pl.set(p, p + 9);
  tr.sequence.assign(in->seq_surfaces + in->seq_offsets[t],
                     in->seq_surfaces + in->seq_offsets[t + 1]);
  const int64_t m0 = in->meas_offsets[t], m1 = in->meas_offsets[t + 1];
  tr.measSurface.assign(in->meas_surfaces + m0, in->meas_surfaces + m1);
  tr.measValue.assign(in->meas_values + 2 * m0, in->meas_values + 2 * m1);
  tr.measCov.assign(in->meas_cov + 4 * m0, in->meas_cov + 4 * m1);
  tr.measDim.clear();
  tr.measIdx.clear();
  if (in->meas_dims) tr.measDim.assign(in->meas_dims + m0, in->meas_dims + m1);
  if (in->meas_indices) tr.measIdx.assign(in->meas_indices + 2 * m0, in->meas_indices + 2 * m1);
}

template <int R, int C>
void store(double* dst, const Mat<R, C>& m) {
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) dst[i * C + j] = m.a[i][j];
}

void storeTrack(const acts_kf_batch_in* in, const acts_kf_batch_out* out,
                int64_t t, const TrackOut& tr) {
  if (out->status) out->status[t] = tr.status;
  if (out->n_states) out->n_states[t] = static_cast<int32_t>(tr.states.size());
  if (out->chi2) out->chi2[t] = tr.chi2;
  if (out->ndf) out->ndf[t] = tr.ndf;
  if (out->n_measurements) out->n_measurements[t] = tr.nMeas;
  if (out->n_holes) out->n_holes[t] = tr.nHoles;
  if (out->n_outliers) out->n_outliers[t] = tr.nOutliers;
  if (out->n_steps) out->n_steps[t] = tr.nSteps;
  if (out->n_step_trials) out->n_step_trials[t] = tr.nTrials;
  if (out->reversed) out->reversed[t] = (!tr.status && tr.reversed) ? 1 : 0;
  const int64_t base = in->seq_offsets[t];
  const int64_t cap = in->seq_offsets[t + 1] - base;
  // slots without a state (past the tip, all of a failed track's) hold an
  // empty slot: surface and measurement -1, flags 0, numbers 0
  const int64_t nDef = tr.status ? 0 : std::min<int64_t>(cap, static_cast<int64_t>(tr.states.size()));
  for (int64_t k = nDef; k < cap; ++k) {
    const int64_t slot = base + k;
    if (out->state_surface) out->state_surface[slot] = -1;
    if (out->state_flags) out->state_flags[slot] = 0u;
    if (out->state_path) out->state_path[slot] = 0.0;
    if (out->state_chi2) out->state_chi2[slot] = 0.f;
    double* arr6[] = {out->predicted, out->filtered, out->smoothed};
    for (double* a : arr6)
      if (a) std::fill(a + 6 * slot, a + 6 * slot + 6, 0.0);
    double* arr36[] = {out->predicted_cov, out->filtered_cov, out->smoothed_cov, out->jacobian};
    for (double* a : arr36)
      if (a) std::fill(a + 36 * slot, a + 36 * slot + 36, 0.0);
    if (out->state_measurement) out->state_measurement[slot] = -1;
  }
  for (size_t k = 0; static_cast<int64_t>(k) < nDef; ++k) {
    const StateOut& s = tr.states[k];
    const int64_t slot = base + static_cast<int64_t>(k);
    if (out->state_surface) out->state_surface[slot] = s.surface;
    if (out->state_flags) out->state_flags[slot] = s.flags;
    if (out->state_path) out->state_path[slot] = s.path;
    if (out->state_chi2) out->state_chi2[slot] = s.chi2;
    if (out->predicted) store(out->predicted + 6 * slot, s.predicted);
    if (out->predicted_cov) store(out->predicted_cov + 36 * slot, s.predictedCov);
    if (out->filtered) store(out->filtered + 6 * slot, s.filtered);
    if (out->filtered_cov) store(out->filtered_cov + 36 * slot, s.filteredCov);
    if (out->smoothed) store(out->smoothed + 6 * slot, s.smoothed);
    if (out->smoothed_cov) store(out->smoothed_cov + 36 * slot, s.smoothedCov);
    if (out->jacobian) store(out->jacobian + 36 * slot, s.jacobian);
    if (out->state_measurement)
      out->state_measurement[slot] =
          s.meas >= 0 ? static_cast<int32_t>(in->meas_offsets[t] + s.meas) : -1;
  }
  if (tr.hasReference) {
    if (out->reference_params) store(out->reference_params + 6 * t, tr.referenceParams);
    if (out->reference_cov) store(out->reference_cov + 36 * t, tr.referenceCov);
  }
}

}  // namespace

extern "C" {

// Fit a batch on the CPU with n_threads worker threads (<= 0: all cores).
// Returns the wall time in seconds through *seconds (may be NULL).
int acts_oracle_fit(const acts_kf_surface* surfaces, int32_t n_surfaces,
                    const acts_kf_material_slab* materials, int32_t n_materials,
                    const acts_kf_field* field, const acts_kf_options* options,
                    const acts_kf_batch_in* in, const acts_kf_batch_out* out,
                    int32_t n_threads, double* seconds) {
  if (!surfaces || !field || !options || !in || !out) return ACTS_KF_ERR_INVALID_ARGUMENT;
  if (!fieldSupported(field)) return ACTS_KF_ERR_UNSUPPORTED;
  Geometry geo;
  buildGeometry(geo, surfaces, n_surfaces, materials, n_materials, field, options);
  if (n_threads <= 0) n_threads = static_cast<int32_t>(std::thread::hardware_concurrency());
  if (n_threads < 1) n_threads = 1;
  std::atomic<int64_t> next{0};
  const int64_t n = in->n_tracks;
  auto t0 = std::chrono::steady_clock::now();
  auto worker = [&]() {
    TrackIn tr;
    TrackOut to;
    while (true) {
      const int64_t t = next.fetch_add(1);
      if (t >= n) break;
      loadTrack(in, t, tr);
      fitTrack(geo, tr, to);
      storeTrack(in, out, t, to);
    }
  };
  if (n_threads == 1) {
    worker();
  } else {
    std::vector<std::thread> pool;
    for (int i = 0; i < n_threads; ++i) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
  }
  auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return ACTS_KF_OK;
}

// The track-state table of a fitted batch (host buffers throughout).
int acts_oracle_track_states(const acts_kf_surface* surfaces, int32_t n_surfaces,
                             const acts_kf_material_slab* materials, int32_t n_materials,
                             const acts_kf_field* field, const acts_kf_options* options,
                             const acts_kf_batch_in* in, const acts_kf_batch_out* fitted,
                             const double* truth_meas, const acts_kf_track_states_out* out) {
  if (!surfaces || !field || !options || !in || !fitted || !out) return ACTS_KF_ERR_INVALID_ARGUMENT;
  if (in->n_tracks <= 0 || out->n_slots <= 0) return ACTS_KF_OK;
  Geometry geo;
  buildGeometry(geo, surfaces, n_surfaces, materials, n_materials, field, options);
  trackStatesTable(geo, *in, *fitted, truth_meas, *out);
  return ACTS_KF_OK;
}

// GainMatrixUpdater for one 2D measurement (loc0, loc1).
int acts_oracle_gain_update(const double* predicted, const double* predicted_cov,
                            const double* meas, const double* meas_cov,
                            int32_t use_joseph, double* filtered,
                            double* filtered_cov, double* chi2) {
  Vec6 x, f;
  M6 P, C;
  for (int i = 0; i < 6; ++i) x.a[i][0] = predicted[i];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) P.a[i][j] = predicted_cov[6 * i + j];
  double c2 = 0;
  int r = gainMatrixUpdate(x, P, meas, meas_cov, use_joseph != 0, f, C, c2);
  if (r) return r;
  store(filtered, f);
  store(filtered_cov, C);
  *chi2 = c2;
  return 0;
}

// The same for a measurement of dimension dim (1 or 2) on the bound
// parameters idx[0..dim-1]: values meas[0..dim-1], covariance meas_cov
// (dim x dim row-major in its first entries).
int acts_oracle_gain_update_sub(const double* predicted, const double* predicted_cov,
                                int32_t dim, const int32_t* idx, const double* meas,
                                const double* meas_cov, int32_t use_joseph,
                                double* filtered, double* filtered_cov, double* chi2) {
  if (dim != 1 && dim != 2) return ACTS_KF_ERR_INVALID_ARGUMENT;
  Vec6 x, f;
  M6 P, C;
  for (int i = 0; i < 6; ++i) x.a[i][0] = predicted[i];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) P.a[i][j] = predicted_cov[6 * i + j];
  Meas m;
  m.dim = dim;
  for (int i = 0; i < dim; ++i) {
    m.idx[i] = idx[i];
    m.v[i] = meas[i];
  }
  for (int i = 0; i < dim * dim; ++i) m.V[i] = meas_cov[i];
  double c2 = 0;
  int r = gainMatrixUpdate(x, P, m, use_joseph != 0, f, C, c2);
  if (r) return r;
  store(filtered, f);
  store(filtered_cov, C);
  *chi2 = c2;
  return 0;
}

// GainMatrixSmoother over n states (arrays of n entries, index 0 first);
// the last state is the tip (smoothed = filtered).
int acts_oracle_smooth(int32_t n, const double* predicted, const double* predicted_cov,
                       const double* filtered, const double* filtered_cov,
                       const double* jacobian, double* smoothed,
                       double* smoothed_cov) {
  std::vector<Vec6> xp(n), xf(n), xs(n);
  std::vector<M6> Cp(n), Cf(n), Cs(n), J(n);
  for (int k = 0; k < n; ++k) {
    for (int i = 0; i < 6; ++i) {
      xp[k].a[i][0] = predicted[6 * k + i];
      xf[k].a[i][0] = filtered[6 * k + i];
      for (int j = 0; j < 6; ++j) {
        Cp[k].a[i][j] = predicted_cov[36 * k + 6 * i + j];
        Cf[k].a[i][j] = filtered_cov[36 * k + 6 * i + j];
        J[k].a[i][j] = jacobian[36 * k + 6 * i + j];
      }
    }
  }
  xs[n - 1] = xf[n - 1];
  Cs[n - 1] = Cf[n - 1];
  for (int k = n - 2; k >= 0; --k) {
    int r = gainMatrixSmoothStep(xf[k], Cf[k], xp[k + 1], Cp[k + 1], xs[k + 1],
                                 Cs[k + 1], J[k + 1], xs[k], Cs[k]);
    if (r) return r;
  }
  for (int k = 0; k < n; ++k) {
    store(smoothed + 6 * k, xs[k]);
    store(smoothed_cov + 36 * k, Cs[k]);
  }
  return 0;
}

// Interactions.cpp functions on a homogeneous slab (mode: 0 Bethe,
// 1 Landau, 2 Landau FWHM, 3 Landau sigma(q/p), 4 theta0).
float acts_oracle_interaction(int32_t mode, const acts_kf_material_slab* m,
                              int32_t abs_pdg, float mass, float qop,
                              float abs_q) {
  SlabF slab;
  if (m) slab = SlabF::make(materialFromMolarDensity(m->x0, m->l0, m->ar, m->z, m->molar_rho),
                            m->thickness);
  switch (mode) {
    case 0: return energyLossBethe(slab, mass, qop, abs_q);
    case 1: return energyLossLandau(slab, mass, qop, abs_q);
    case 2: return energyLossLandauFwhm(slab, mass, qop, abs_q);
    case 3: return energyLossLandauSigmaQOverP(slab, mass, qop, abs_q);
    case 4: return multipleScatteringTheta0(slab, abs_pdg, mass, qop, abs_q);
    default: return 0.f;
  }
}

// Surface helpers for tests: intersection (returns number of intersections,
// paths[2], status[2]), global<->local, jacobians.
int acts_oracle_surface_intersect(const acts_kf_surface* s, const double* pos,
                                  const double* dir, double tol, double* paths,
                                  int32_t* status) {
  Surface sf;
  sf.type = s->type;
  sf.pl.set(s->rotation, s->translation);
  sf.b0 = s->bounds[0];
  sf.b1 = s->bounds[1];
  MultiIntersection mi = sf.intersect(v3(pos[0], pos[1], pos[2]),
                                      v3(dir[0], dir[1], dir[2]), tol);
  for (int i = 0; i < 2; ++i) {
    paths[i] = mi.is[i].path;
    status[i] = static_cast<int32_t>(mi.is[i].status);
  }
  return mi.size;
}

int acts_oracle_surface_local(const acts_kf_surface* s, int32_t to_global,
                              const double* in3, double* out3) {
  Surface sf;
  sf.type = s->type;
  sf.pl.set(s->rotation, s->translation);
  sf.b0 = s->bounds[0];
  sf.b1 = s->bounds[1];
  if (to_global) {
    V3 g = sf.localToGlobal(in3[0], in3[1]);
    for (int i = 0; i < 3; ++i) out3[i] = g[i];
    return 0;
  }
  double l0, l1;
  if (!sf.globalToLocal(v3(in3[0], in3[1], in3[2]), l0, l1))
    return ACTS_KF_STATUS_NOT_ON_SURFACE;
  out3[0] = l0;
  out3[1] = l1;
  out3[2] = 0;
  return 0;
}

// LineSurface local <-> global with the direction the frame depends on.
int acts_oracle_line_local(const acts_kf_surface* s, int32_t to_global, const double* in3,
                           const double* dir, double* out3) {
  Surface sf;
  sf.type = s->type;
  sf.pl.set(s->rotation, s->translation);
  const V3 d = v3(dir[0], dir[1], dir[2]);
  if (to_global) {
    V3 g = sf.localToGlobal(in3[0], in3[1], d);
    for (int i = 0; i < 3; ++i) out3[i] = g[i];
    return 0;
  }
  double l0, l1;
  if (!sf.globalToLocal(v3(in3[0], in3[1], in3[2]), d, l0, l1))
    return ACTS_KF_STATUS_NOT_ON_SURFACE;
  out3[0] = l0;
  out3[1] = l1;
  out3[2] = 0;
  return 0;
}

// Propagate bound parameters (and optionally covariance) from a start plane
// to a target surface; returns status, end params/cov/jacobian, path, steps.
int acts_oracle_propagate(const acts_kf_field* field, const acts_kf_options* options,
                          const double* start_plane, const double* start_params,
                          const double* start_cov, const acts_kf_surface* target,
                          double direction, double* end_params, double* end_cov,
                          double* jacobian, double* path, int32_t* n_steps) {
  Geometry geo;
  buildGeometry(geo, target, 0, nullptr, 0, field, options);
  Surface start;
  start.type = ACTS_KF_SURFACE_PLANE;
  start.pl.set(start_plane, start_plane + 9);
  Surface tgt;
  tgt.type = target->type;
  tgt.pl.set(target->rotation, target->translation);
  tgt.b0 = target->bounds[0];
  tgt.b1 = target->bounds[1];
  Vec6 sp;
  for (int i = 0; i < 6; ++i) sp.a[i][0] = start_params[i];
  M6 sc;
  if (start_cov)
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) sc.a[i][j] = start_cov[6 * i + j];
  Vec6 ep;
  M6 ec = M6::zero(), jac = M6::identity();
  double pl = 0;
  int ns = 0;
  int r = propagateToSurface(geo, sp, start_cov ? &sc : nullptr, start, tgt,
                             direction, ep, ec, jac, pl, ns);
  if (r) return r;
  store(end_params, ep);
  if (end_cov) store(end_cov, ec);
  if (jacobian) store(jacobian, jac);
  if (path) *path = pl;
  if (n_steps) *n_steps = ns;
  return 0;
}

// Field lookups: the oracle's InterpolatedBFieldMap / ConstantBField.
int acts_oracle_field_eval(const acts_kf_field* field, int64_t n, const double* pos,
                           double* out, int32_t* status) {
  if (!field || !pos || !out || !status || !fieldSupported(field))
    return ACTS_KF_ERR_INVALID_ARGUMENT;
  Field f;
  f.kind = field->kind;
  for (int i = 0; i < 3; ++i) f.B[i] = field->constant[i];
  if (field->kind == ACTS_KF_FIELD_GRID_XYZ)
    f.grid.set(field->n, field->point_min, field->point_max, field->values);
  for (int64_t i = 0; i < n; ++i) {
    V3 B = v3(0, 0, 0);
    status[i] = f.get(v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]), B);
    for (int k = 0; k < 3; ++k) out[3 * i + k] = B[k];
  }
  return 0;
}

// SolenoidBField::getField at n positions.
int acts_oracle_solenoid_eval(double radius, double length, int32_t nCoils, double bCenter,
                              int64_t n, const double* pos, double* out) {
  if (!pos || !out || nCoils <= 0) return ACTS_KF_ERR_INVALID_ARGUMENT;
  const Solenoid s(radius, length, nCoils, bCenter);
  for (int64_t i = 0; i < n; ++i) {
    const V3 B = s.field(v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
    for (int k = 0; k < 3; ++k) out[3 * i + k] = B[k];
  }
  return 0;
}

// Complete elliptic integrals K, E of modulus k (what SolenoidBField calls
// ellint_1/ellint_2 with).
void acts_oracle_elliptic(double k, double* K, double* E) { ellipticKE(k, *K, *E); }

}  // extern "C"
