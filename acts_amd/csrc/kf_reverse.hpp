// kf_reverse.hpp -- reverse filtering (KalmanFitter.hpp:884-945).
//
// After the forward filter, a track that the reverse-filtering decision
// picks is filtered again backwards: from its last measurement state's
// filtered parameters, covariance scaled by reverseFilteringCovarianceScaling,
// on that state's surface, through the same surface sequence in the other
// direction (DirectNavigator.hpp:222-272), with the reference surface as
// the actor's target (make_propagator_options, KalmanFitter.hpp:735-792).
// The pass runs in the persistent forward kernel (Rev = true: backward
// steps and navigation, material split by direction, no state records).
// Then the first measurement state's smoothed parameters are the reverse
// filter's last measurement state's filtered ones, the last measurement
// state's smoothed parameters its own filtered ones, and no other state has
// smoothed parameters.
#pragma once

#include "kf_reference.hpp"

namespace kf {

// Marker in DevScratch::track (first measurement row) of a track whose
// reverse filter reached the reference surface: its parameters there are
// the fit's, and the extrapolation kernel leaves it alone.
constexpr int32_t kReferenceSet = -3;

// Whether the track in queue slot q is filtered in reverse, after its
// forward pass: KalmanFitterOptions::reverseFiltering, or the Examples'
// SimpleReverseFilteringLogic::doBackwardFiltering (KalmanFitterFunction.cpp:
// 53-57) on the last measurement state: |1 / filtered q/p| <= threshold.
__device__ inline bool reverseWanted(const RevArgs& rv, const DevScratch& sc, int64_t q) {
  if (sc.track[q] != 0) return false;
  const int lastMeas = sc.track[3 * sc.stride + q];
  if (lastMeas < 0) return false;
  if (rv.all) return true;
  return fabs(1.0 / sd(sc, lastMeas, kFilt + 4, q)) <= rv.threshold;
}

// fit_impl's reverse start (KalmanFitter.hpp:904-911) and filter_impl's
// set-up: EigenStepper::initialize from the last measurement state's
// filtered parameters and scaled covariance on its surface, the
// DirectNavigator next to the surface's first entry of the sequence (or at
// the sequence's start going backwards), the loop protection; the actor's
// first act is on the start surface, which is the navigator's current
// surface (so the lane starts at the surface update).  Tracks that are not
// filtered in reverse leave the lane idle.
template <bool Mixed, bool Sympy = false>
__device__ inline int laneInitReverse(Lane& L, const int64_t q, const int64_t t,
                                      const FwdArgs& a, const DevSurface* geo) {
  const DevScratch& sc = a.sc;
  if (!reverseWanted(a.rv, sc, q)) return kPhaseIdle;
  const DevBatch& in = a.in;
  const DevOpts& o = a.o;
  L.t = t;
  L.q = q;
  const int64_t s0 = in.seqOff[t];
  const int seqLen = static_cast<int>(in.seqOff[t + 1] - s0);
  L.m0 = in.measOff[t];
  L.nMeas = static_cast<int>(in.measOff[t + 1] - L.m0);
  L.seq.p = in.seqSurf;
  L.seq.s0 = s0;
  L.status = 0;
  L.nMeasStates = 0;
  L.nStates = 0;
  L.tip = -1;
  L.firstMeas = -1;
  L.lastMeas = -1;
  L.iter = 0;
  L.finished = false;
  L.lastSurfIter = 0;
  L.noTransport = false;
  L.replaying = false;
  L.noSkip = false;
  L.nUnique = uniqueMeasurementSurfaces(in, L.m0, L.nMeas);
  L.mNext = 0;
  a.rv.reached[q] = 0;
  const int lastMeas = sc.track[3 * sc.stride + q];
  const int surf = si(sc, lastMeas, kISurf, q);
  double p[27];
#pragma unroll
  for (int i = 0; i < 6; ++i) p[i] = sd(sc, lastMeas, kFilt + i, q);
#pragma unroll
  for (int i = 0; i < 21; ++i) p[6 + i] = sd(sc, lastMeas, kFCov + i, q) * a.rv.scale;
  Track& T = L.T;
  initializeFromBound<Sympy>(T, geo[surf], p, p + 6, o);
  L.startSurf = surf;
  int found = -1;
  if (o.navigator == ACTS_KF_NAVIGATOR_DIRECT)
    for (int i = 0; i < seqLen; ++i)
      if (found < 0 && L.seq(i) == surf) found = i;
  // the layer navigator resolves its lists without the start surface (the
  // volume's layer at the start position is the start object)
  L.nv.seqPos = o.navigator == ACTS_KF_NAVIGATOR_DIRECT ? (found >= 0 ? found + 1 : seqLen)
                                                       : kNavUnresolved;
  L.nv.seqLen = seqLen;
  L.nv.navBreak = false;
  L.nv.tSurf = -1;
  L.nv.tIdx = 0;
  if (o.navigator != ACTS_KF_NAVIGATOR_DIRECT) {
    clearReached(layerCtx(a, q));
    markReached(layerCtx(a, q), surf);
  }
  L.propLimit = o.pathLimit;
  loopProtection<Mixed>(T, a.F, o, L.propLimit, false);
  L.kfLimit = DBL_MAX;
  L.reached = surf;
  return kPhaseSurface;
}

// SurfaceReached::checkAbort (StandardAborters.hpp:94-155) with the
// reference surface as target, going backwards: reached if the line's
// intersection along the propagation direction is on the surface; the
// actor step size is set to the intersection (released first,
// updateSingleSurfaceStatus, SteppingHelper.hpp:41-85) unless it lies
// behind the near limit (-100 um).  When reached, the bound state on the
// line is the filter's fitted parameters (KalmanFitter.hpp:404-418).
template <bool Sympy>
__device__ inline bool reverseTargetReached(Lane& L, const FwdArgs& a) {
  Track& T = L.T;
  const V3 d = mk(-T.dir.x, -T.dir.y, -T.dir.z);
  double path;
  const int st = lineIntersect(a.rv.target, T.pos, d, a.o.surfaceTolerance, path);
  if (st != kUnreachable && path > -100 * 1e-3 && path < DBL_MAX + kOnSurfaceTolerance) {
    T.cActor = kNotSet;
    updateStep(T.cActor, path);
  }
  if (st != kOnSurface) return false;
  double* r = a.rv.rec + kRevRecord * L.q + kRevTarget;
  const int bs = lineBoundState<Sympy>(T, a.rv.target, kOnSurfaceTolerance, r, r + 6);
  if (bs) {
    L.status = bs;
  } else {
    a.rv.reached[L.q] = 1;
  }
  return true;
}

// A fit that fails after the forward pass: status, empty summary, every slot
// an empty slot (as a failed extrapolation leaves it).
__device__ inline void failFit(const DevOut& out, const DevBatch& in, const DevScratch& sc,
                               int64_t q, int64_t t, int status) {
  sc.track[q] = status;
  for (int64_t slot = in.seqOff[t]; slot < in.seqOff[t + 1]; ++slot) clearSlot(out, slot);
  if (out.status) out.status[t] = status;
  if (out.nStates) out.nStates[t] = 0;
  if (out.chi2) out.chi2[t] = 0.f;
  if (out.ndf) out.ndf[t] = 0;
  if (out.nMeas) out.nMeas[t] = 0;
  if (out.nHoles) out.nHoles[t] = 0;
  if (out.nOutliers) out.nOutliers[t] = 0;
  if (out.reversed) out.reversed[t] = 0;
}

// Write a packed covariance as the row-major 6x6 output of slot `slot`.
__device__ inline void storeSymSlot(double* dst, const double* p) {
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) dst[6 * i + j] = p[sym(i, j)];
}

// The end of a reverse pass (KalmanFitter.hpp:913-945): a failed reverse
// filter fails the fit; its last measurement state must lie on the forward
// track's first measurement surface (else InconsistentTrackStates); the
// first measurement state's smoothed parameters become the reverse filter's
// filtered ones, the last measurement state's its own filtered ones (the
// smoother kernel left every other smoothed slot NaN); the reference-surface
// states for the extrapolation follow, unless the reverse filter reached
// the reference surface itself.
__device__ inline void reverseFinish(Lane& L, const FwdArgs& a) {
  const DevScratch& sc = a.sc;
  const DevOut& out = a.out;
  const DevBatch& in = a.in;
  const int64_t q = L.q, t = L.t;
  int status = L.status;
  if (!status && L.nMeasStates == 0) status = ACTS_KF_STATUS_NO_MEASUREMENT_FOUND;
  const int firstMeas = sc.track[2 * sc.stride + q];
  const int lastMeas = sc.track[3 * sc.stride + q];
  const double* r = a.rv.rec + kRevRecord * q;
  if (!status && static_cast<int>(r[27]) != si(sc, firstMeas, kISurf, q))
    status = ACTS_KF_STATUS_INCONSISTENT_TRACK_STATES;
  if (status) {
    failFit(out, in, sc, q, t, status);
    return;
  }
  const int64_t base = in.seqOff[t];
  double lf[27];
#pragma unroll
  for (int i = 0; i < 6; ++i) lf[i] = sd(sc, lastMeas, kFilt + i, q);
#pragma unroll
  for (int i = 0; i < 21; ++i) lf[6 + i] = sd(sc, lastMeas, kFCov + i, q);
  const int64_t slotLast = base + lastMeas, slotFirst = base + firstMeas;
  if (out.smoothed)
    for (int i = 0; i < 6; ++i) out.smoothed[6 * slotLast + i] = lf[i];
  if (out.smoothedCov) storeSymSlot(out.smoothedCov + 36 * slotLast, lf + 6);
  if (out.smoothed)
    for (int i = 0; i < 6; ++i) out.smoothed[6 * slotFirst + i] = r[i];
  if (out.smoothedCov) storeSymSlot(out.smoothedCov + 36 * slotFirst, r + 6);
  if (out.reversed) out.reversed[t] = 1;
  if (!sc.ref) return;
  if (a.rv.reached[q]) {
    // the fitted parameters on the reference surface come from the reverse
    // filter; the extrapolation is skipped
    if (out.referenceParams)
      for (int i = 0; i < 6; ++i) out.referenceParams[6 * t + i] = r[kRevTarget + i];
    if (out.referenceCov)
      for (int i = 0; i < 36; ++i) out.referenceCov[36 * t + i] = r[kRevTarget + 6 + i];
    sc.track[2 * sc.stride + q] = kReferenceSet;
    return;
  }
  double* rf = sc.ref + kRefComps * q;
#pragma unroll
  for (int i = 0; i < 27; ++i) rf[i] = r[i];
#pragma unroll
  for (int i = 0; i < 27; ++i) rf[27 + i] = lf[i];
  if (firstMeas == lastMeas) {
#pragma unroll
    for (int i = 0; i < 27; ++i) rf[27 + i] = r[i];
  }
}

}  // namespace kf
