// acts_kf.hpp -- C++ facade over the C ABI of acts_kf.h, shaped like the
// reference call
//
//   Acts::KalmanFitter<Propagator<EigenStepper<>, DirectNavigator>, traj_t>
//     ::fit(begin, end, startParameters, kfOptions, surfaceSequence, tracks)
//   (Core/include/Acts/TrackFitting/KalmanFitter.hpp:721-731), and, with
//   options.navigator = ACTS_KF_NAVIGATOR_LAYERS, the Navigator overload
//     ::fit(begin, end, startParameters, kfOptions, tracks)
//   (KalmanFitter.hpp:693-699), which finds the surfaces from the geometry
//
// Header-only; depends on nothing but the C ABI.  Two ways to call it:
//   - fit(...): one track, synchronous, returns a Result like the reference;
//   - enqueue(...) for many tracks, then fitBatch(): the batched path the GPU
//     is built for (one kernel launch pair for the whole batch).
#ifndef ACTS_KF_HPP
#define ACTS_KF_HPP

#include <algorithm>
#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "acts_kf.h"

namespace acts_amd {

// BoundTrackParameters on a plane (rotation columns = local axes).
struct BoundTrackParameters {
  std::array<double, 6> params{};        // loc0, loc1, phi, theta, q/p, t
  std::array<double, 36> covariance{};   // row-major
  std::array<double, 9> planeRotation{1, 0, 0, 0, 1, 0, 0, 0, 1};
  std::array<double, 3> planeCenter{};
};

// A 2D measurement (source link + calibrated values) on a geometry surface.
struct Measurement {
  int32_t surface = -1;
  std::array<double, 2> values{};
  std::array<double, 4> covariance{};    // row-major 2x2
};

// One fitted track state (the track-state table row of the reference).
struct TrackState {
  int32_t surface = -1;
  uint32_t flags = 0;                    // ACTS_KF_STATE_*
  double pathLength = 0;
  float chi2 = 0;
  std::array<double, 6> smoothed{};
  std::array<double, 36> smoothedCovariance{};
};

// Result<TrackProxy> of the reference: status 0 = ok, otherwise the
// reference's error code (category | enum value).
struct FitResult {
  int32_t status = 0;
  float chi2 = 0;
  int32_t ndf = 0, nMeasurements = 0, nHoles = 0, nOutliers = 0;
  // filtered in reverse instead of smoothed (reverse filtering options)
  bool reversed = false;
  std::vector<TrackState> states;
  // with a reference surface in the options: the track parameters and
  // covariance on it (TrackProxy::parameters()/covariance())
  bool hasReference = false;
  std::array<double, 6> referenceParameters{};
  std::array<double, 36> referenceCovariance{};
  bool ok() const { return status == ACTS_KF_STATUS_OK; }
  std::string message() const { return acts_kf_status_string(status); }
};

class KalmanFitter {
 public:
  KalmanFitter(const std::vector<acts_kf_surface>& surfaces,
               const std::vector<acts_kf_material_slab>& materials,
               const acts_kf_field& field, const acts_kf_options& options, int device = 0)
      : withReference_(options.reference_surface != ACTS_KF_REFERENCE_NONE),
        nSurfaces_(surfaces.size()) {
    const int rc = acts_kf_create(surfaces.data(), static_cast<int32_t>(surfaces.size()),
                                  materials.data(), static_cast<int32_t>(materials.size()),
                                  &field, &options, device, &ctx_);
    if (rc != ACTS_KF_OK) throw std::runtime_error(std::string("acts_kf_create: ") +
                                                   acts_kf_last_error());
  }
  KalmanFitter(const KalmanFitter&) = delete;
  KalmanFitter& operator=(const KalmanFitter&) = delete;
  ~KalmanFitter() {
    if (ctx_) acts_kf_destroy(ctx_);
  }

  static acts_kf_options defaultOptions() {
    acts_kf_options o;
    acts_kf_default_options(&o);
    return o;
  }

  // Queue one track (measurements [begin, end), start parameters, surface
  // sequence); returns its index in the next fitBatch() result.
  template <typename MeasIt>
  size_t enqueue(MeasIt begin, MeasIt end, const BoundTrackParameters& start,
                 const std::vector<int32_t>& surfaceSequence) {
    const size_t idx = seqOffsets_.size() - 1;
    params_.insert(params_.end(), start.params.begin(), start.params.end());
    cov_.insert(cov_.end(), start.covariance.begin(), start.covariance.end());
    plane_.insert(plane_.end(), start.planeRotation.begin(), start.planeRotation.end());
    plane_.insert(plane_.end(), start.planeCenter.begin(), start.planeCenter.end());
    seqSurfaces_.insert(seqSurfaces_.end(), surfaceSequence.begin(), surfaceSequence.end());
    seqOffsets_.push_back(static_cast<int64_t>(seqSurfaces_.size()));
    for (MeasIt it = begin; it != end; ++it) {
      const Measurement& m = *it;
      measSurfaces_.push_back(m.surface);
      measValues_.insert(measValues_.end(), m.values.begin(), m.values.end());
      measCov_.insert(measCov_.end(), m.covariance.begin(), m.covariance.end());
    }
    measOffsets_.push_back(static_cast<int64_t>(measSurfaces_.size()));
    maxLen_ = std::max<int32_t>(maxLen_, static_cast<int32_t>(surfaceSequence.size()));
    return idx;
  }

  // The same for the layer navigator: no surface sequence, `stateSlots`
  // states at most (the track fails with ACTS_KF_STATUS_STATE_CAPACITY if
  // the navigator finds more).
  template <typename MeasIt>
  size_t enqueue(MeasIt begin, MeasIt end, const BoundTrackParameters& start,
                 size_t stateSlots) {
    return enqueue(begin, end, start, std::vector<int32_t>(stateSlots, -1));
  }

  size_t queued() const { return seqOffsets_.size() - 1; }

  // Fit every queued track (host buffers, one batch) and clear the queue.
  std::vector<FitResult> fitBatch() {
    const int64_t n = static_cast<int64_t>(queued());
    const int64_t slots = seqOffsets_.back();
    std::vector<FitResult> results(static_cast<size_t>(n));
    if (n == 0) return results;
    std::vector<int32_t> status(n), nStates(n), ndf(n), nMeas(n), nHoles(n), nOut(n),
        reversed(n);
    std::vector<float> chi2(n), stateChi2(slots);
    std::vector<int32_t> stateSurface(slots);
    std::vector<uint32_t> stateFlags(slots);
    std::vector<double> statePath(slots), smoothed(6 * slots), smoothedCov(36 * slots);
    std::vector<double> refParams(withReference_ ? 6 * n : 0),
        refCov(withReference_ ? 36 * n : 0);
    acts_kf_batch_in in{};
    in.n_tracks = n;
    in.start_params = params_.data();
    in.start_cov = cov_.data();
    in.start_plane = plane_.data();
    in.seq_offsets = seqOffsets_.data();
    in.seq_surfaces = seqSurfaces_.data();
    in.meas_offsets = measOffsets_.data();
    in.meas_surfaces = measSurfaces_.data();
    in.meas_values = measValues_.data();
    in.meas_cov = measCov_.data();
    in.max_sequence_length = maxLen_;
    acts_kf_batch_out out{};
    out.status = status.data();
    out.n_states = nStates.data();
    out.chi2 = chi2.data();
    out.ndf = ndf.data();
    out.n_measurements = nMeas.data();
    out.n_holes = nHoles.data();
    out.n_outliers = nOut.data();
    out.reversed = reversed.data();
    out.state_surface = stateSurface.data();
    out.state_flags = stateFlags.data();
    out.state_path = statePath.data();
    out.state_chi2 = stateChi2.data();
    out.smoothed = smoothed.data();
    out.smoothed_cov = smoothedCov.data();
    if (withReference_) {
      out.reference_params = refParams.data();
      out.reference_cov = refCov.data();
    }
    check(acts_kf_fit_batch(ctx_, &in, &out, nullptr), "acts_kf_fit_batch");
    check(acts_kf_sync(ctx_), "acts_kf_sync");
    for (int64_t t = 0; t < n; ++t) {
      FitResult& r = results[static_cast<size_t>(t)];
      r.status = status[t];
      r.chi2 = chi2[t];
      r.ndf = ndf[t];
      r.nMeasurements = nMeas[t];
      r.nHoles = nHoles[t];
      r.nOutliers = nOut[t];
      r.reversed = reversed[t] != 0;
      if (!r.ok()) continue;
      if (withReference_) {
        r.hasReference = true;
        for (int i = 0; i < 6; ++i) r.referenceParameters[i] = refParams[6 * t + i];
        for (int i = 0; i < 36; ++i) r.referenceCovariance[i] = refCov[36 * t + i];
      }
      r.states.resize(static_cast<size_t>(nStates[t]));
      for (int32_t k = 0; k < nStates[t]; ++k) {
        const int64_t s = seqOffsets_[t] + k;
        TrackState& st = r.states[static_cast<size_t>(k)];
        st.surface = stateSurface[s];
        st.flags = stateFlags[s];
        st.pathLength = statePath[s];
        st.chi2 = stateChi2[s];
        for (int i = 0; i < 6; ++i) st.smoothed[i] = smoothed[6 * s + i];
        for (int i = 0; i < 36; ++i) st.smoothedCovariance[i] = smoothedCov[36 * s + i];
      }
    }
    clear();
    return results;
  }

  // The reference call shape: one track, synchronous.
  template <typename MeasIt>
  FitResult fit(MeasIt begin, MeasIt end, const BoundTrackParameters& start,
                const std::vector<int32_t>& surfaceSequence) {
    if (queued() != 0) throw std::logic_error("fit() with tracks queued: call fitBatch()");
    enqueue(begin, end, start, surfaceSequence);
    return std::move(fitBatch().front());
  }

  // The Navigator overload (layer navigator): as many state slots as the
  // geometry has surfaces, unless given.
  template <typename MeasIt>
  FitResult fit(MeasIt begin, MeasIt end, const BoundTrackParameters& start,
                size_t stateSlots = 0) {
    if (queued() != 0) throw std::logic_error("fit() with tracks queued: call fitBatch()");
    enqueue(begin, end, start, stateSlots ? stateSlots : nSurfaces_);
    return std::move(fitBatch().front());
  }

  acts_kf_ctx* context() const { return ctx_; }

 private:
  static void check(int rc, const char* what) {
    if (rc != ACTS_KF_OK)
      throw std::runtime_error(std::string(what) + ": " + acts_kf_last_error());
  }
  void clear() {
    params_.clear();
    cov_.clear();
    plane_.clear();
    seqOffsets_.assign(1, 0);
    seqSurfaces_.clear();
    measOffsets_.assign(1, 0);
    measSurfaces_.clear();
    measValues_.clear();
    measCov_.clear();
    maxLen_ = 0;
  }

  acts_kf_ctx* ctx_ = nullptr;
  bool withReference_ = false;
  size_t nSurfaces_ = 0;
  std::vector<double> params_, cov_, plane_;
  std::vector<int64_t> seqOffsets_{0};
  std::vector<int32_t> seqSurfaces_;
  std::vector<int64_t> measOffsets_{0};
  std::vector<int32_t> measSurfaces_;
  std::vector<double> measValues_, measCov_;
  int32_t maxLen_ = 0;
};

}  // namespace acts_amd

#endif  // ACTS_KF_HPP
