// kf_kernel.hpp -- the fit kernel: one lane per track.
//
// Forward pass: Propagator::propagate with the KalmanFitter actor
// (Propagator.ipp:25-210, KalmanFitter.hpp:339-671), cut into phases (init,
// step, surface update, finish) that a persistent kernel schedules per
// wavefront: lanes that reach their next surface wait until enough lanes of
// the wavefront are waiting, then the (expensive) surface update runs for
// all of them together, and a lane whose track is done takes the next track
// from a global queue.  The order of operations seen by each track is
// unchanged.
// Backward pass (a separate kernel, so that each pass gets its own register
// allocation): GainMatrixSmoother from the tip (GainMatrixSmoother.hpp:43-99)
// and calculateTrackQuantities (TrackHelpers.hpp:421-445).
#pragma once

#include "kf_fit.hpp"

namespace kf {

// Component c of the scratch record of state k of the track in queue slot
// q: component-major with the slots innermost, so that the lanes of a
// wavefront, which hold consecutive slots, store and load one component of
// the same state index as one contiguous run.
__device__ inline double& sd(const DevScratch& sc, int k, int c, int64_t q) {
  return sc.d[(static_cast<int64_t>(k) * kScratchComps + c) * sc.stride + q];
}
__device__ inline int32_t& si(const DevScratch& sc, int k, int c, int64_t q) {
  return sc.i[(static_cast<int64_t>(k) * kIScratchComps + c) * sc.stride + q];
}

// Phases of a lane of the persistent forward kernel.
enum LanePhase : int {
  kPhaseIdle = 0,     // needs a track
  kPhaseStepping = 1, // propagating towards its target surface
  kPhaseSurface = 2,  // reached its target surface, filter pending
  kPhaseFinished = 3, // forward pass over, summary pending
  kPhaseRetired = 4,  // no tracks left
};

struct SeqAt {
  const int32_t* p;
  int64_t s0;
  __device__ int operator()(int i) const { return p[s0 + i]; }
};

// Everything one track needs between two phases of the forward pass.
struct Lane {
  int64_t t;  // track index in the batch
  int64_t q;  // queue slot (index of the track's scratch records)
  int64_t m0;
  int nMeas, nUnique;
  int mNext;  // measurement after the last one matched (all surfaces distinct)
  SeqAt seq;
  Track T;
  Nav nv;
  double propLimit, kfLimit;
  int nMeasStates, nStates, tip;
  int firstMeas, lastMeas;  // states updated with a measurement (not outliers)
  bool finished;
  int status;
  uint32_t iter;
  int reached;
  // steps without the jacobian transport (see FwdArgs::skipAfter)
  uint32_t lastSurfIter;  // step count at the last surface
  bool noTransport;       // stepping without transport since a checkpoint
  bool replaying;         // back at the checkpoint: steps with transport
  bool noSkip;            // replayed once: keeps the transport from now on
  int startSurf;          // reverse filter: its start surface (else -1)
};

// Reverse filtering (KalmanFitter.hpp:884-945): the second pass of the
// forward kernel, backwards from the last measurement state.  `rec` holds
// per queue slot the reverse filter's last measurement state (kRevComps:
// filtered parameters, packed covariance, surface) and, if it reached its
// target, the bound parameters and packed covariance there.
constexpr int kRevComps = 28, kRevTarget = 28, kRevRecord = 70;
struct RevArgs {
  double scale;      // reverseFilteringCovarianceScaling
  double threshold;  // the reverse-filtering logic's momentum threshold
  int32_t all;       // KalmanFitterOptions::reverseFiltering
  int32_t hasTarget; // the reference surface is the reverse filter's target
  LineRef target;
  double* rec;       // [slot][kRevRecord]
  int32_t* reached;  // [slot] 1 if the reverse filter reached its target
};

struct FwdArgs {
  DevBatch in;
  DevOut out;
  DevScratch sc;
  const DevSurface* __restrict__ geo;
  const DevSlab* __restrict__ slabs;
  DevField F;
  DevOpts o;
  PhysConst pc;
  int32_t nSurfaces;
  int32_t nSlabs;
  // A track that has taken skipAfter steps since its last surface copies its
  // lane to ckpt[lane id] and steps on without the transport; if it then
  // reaches a surface, the lane is restored and the steps are repeated with
  // the transport (same trajectory, same results).  0 = never.
  Lane* ckpt;
  int32_t skipAfter;
  int32_t pad;
  RevArgs rv;
};

// The layer-navigator record of queue slot q.
__device__ inline LayerCtx layerCtx(const FwdArgs& a, int64_t q) {
  return LayerCtx{a.sc.nav ? a.sc.nav + q : nullptr, a.sc.nav ? a.sc.navMask + q : nullptr,
                  a.sc.stride, a.nSurfaces};
}

// The reverse filter's actor target check (SurfaceReached::checkAbort on the
// reference surface) and the end of a reverse pass: kf_reverse.hpp.
template <bool Sympy>
__device__ inline bool reverseTargetReached(Lane& L, const FwdArgs& a);
__device__ inline void reverseFinish(Lane& L, const FwdArgs& a);

// The number of distinct measurement surfaces of a track
// (KalmanFitter.hpp:750-755, first one wins): a set of the surface indices
// below 256 in four 64-bit words; pairwise comparison only for a track with
// a measurement on another index.
__device__ inline int uniqueMeasurementSurfaces(const DevBatch& in, int64_t m0, int nMeas) {
  int nUnique = 0;
  bool wide = false;
  uint64_t seen[4] = {0, 0, 0, 0};
  for (int i = 0; i < nMeas; ++i) {
    const int sm = in.measSurf[m0 + i];
    wide = wide || sm < 0 || sm >= 256;
    const uint64_t bit = 1ull << (sm & 63);
    const int w = (sm >> 6) & 3;
    bool dup = false;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      dup = dup || (w == k && (seen[k] & bit) != 0);
      seen[k] |= (w == k) ? bit : 0ull;
    }
    nUnique += dup ? 0 : 1;
  }
  if (wide) {
    nUnique = 0;
    for (int i = 0; i < nMeas; ++i) {
      const int si_ = in.measSurf[m0 + i];
      bool dup = false;
      for (int j = 0; j < i; ++j) {
        const int sj = in.measSurf[m0 + j];
        dup = dup || sj == si_;
      }
      nUnique += dup ? 0 : 1;
    }
  }
  return nUnique;
}

// EigenStepper::initialize (EigenStepper.ipp:37-70), DirectNavigator
// initialisation and the pre-propagation act of the KalmanFitter actor.
// Returns the lane's next phase.
template <bool Mixed, bool Sympy = false>
__device__ inline int laneInit(Lane& L, const int64_t q, const int64_t t, const FwdArgs& a,
                               const DevSurface* geo) {
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
  L.reached = -1;
  L.finished = false;
  L.lastSurfIter = 0;
  L.noTransport = false;
  L.replaying = false;
  L.noSkip = false;
  L.startSurf = -1;
  L.T.nSteps = 0;
  L.T.nTrials = 0;
  // surface indices come from the caller: reject the track if any is out of
  // range instead of reading outside the geometry table
  bool badIndex = false;
  if (o.navigator == ACTS_KF_NAVIGATOR_DIRECT) {
    for (int i = 0; i < seqLen; ++i) {
      const int sid = L.seq(i);
      badIndex = badIndex || sid < 0 || sid >= a.nSurfaces;
    }
  }
  // measurement subspaces the updater does not handle fail the track too
  if (in.measDims || in.measIdx) {
    for (int i = 0; i < L.nMeas; ++i) {
      MeasDev M;
      badIndex = badIndex || !loadMeas(in, L.m0 + i, M);
    }
  }
  if (badIndex) {
    L.status = ACTS_KF_STATUS_PROPAGATION_FAILURE;
    return kPhaseFinished;
  }
  const int nUnique = uniqueMeasurementSurfaces(in, L.m0, L.nMeas);
  L.nUnique = nUnique;
  L.mNext = 0;

  Track& T = L.T;
  {
    const double* bp = in.startParams + 6 * t;
    const double* pl = in.startPlane + 12 * t;
    const double sinTheta = sin(bp[3]);
    T.dir = mk(cos(bp[2]) * sinTheta, sin(bp[2]) * sinTheta, cos(bp[3]));
    T.pos = toGlobal2(pl, pl + 9, bp[0], bp[1]);
    T.time = bp[5];
    T.qop = bp[4];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = i; j < 6; ++j) T.cov[sym(i, j)] = in.startCov[36 * t + 6 * i + j];
    T.jp[0] = pl[0]; T.jp[1] = pl[1];
    T.jp[2] = pl[3]; T.jp[3] = pl[4];
    T.jp[4] = pl[6]; T.jp[5] = pl[7];
    directionB2F(T.dir, T.jd);
    resetTransport(T);
    if constexpr (Sympy) sympyResetM(T);
    T.pathAcc = 0;
    T.cNav = kNotSet;
    T.cActor = kNotSet;
    T.cUser = o.maxStepSize;
    T.cAcc = o.initialStepSize;
  }
  L.nv.seqPos = o.navigator == ACTS_KF_NAVIGATOR_DIRECT ? -1 : kNavUnresolved;
  L.nv.seqLen = seqLen;
  L.nv.navBreak = false;
  L.nv.tSurf = -1;
  L.nv.tIdx = 0;
  if (o.navigator != ACTS_KF_NAVIGATOR_DIRECT) clearReached(layerCtx(a, q));

  L.propLimit = o.pathLimit;
  loopProtection<Mixed>(T, a.F, o, L.propLimit, false);
  L.kfLimit = DBL_MAX;
  // pre-propagation act (the start surface carries neither material nor a
  // measurement, so only the finalisation checks run)
  loopProtection<Mixed>(T, a.F, o, L.kfLimit, true);
  L.finished = (L.nMeasStates == L.nUnique) || pathAbort(T, L.kfLimit, o.surfaceTolerance);
  if (L.finished || pathAbort(T, L.propLimit, o.surfaceTolerance)) return kPhaseFinished;
  L.status = getNextTarget(L.nv, T, geo, o, L.seq, layerCtx(a, q), -1);
  return L.status ? kPhaseFinished : kPhaseStepping;
}

template <bool Mixed, bool Sympy = false, bool Rev = false>
__device__ inline int laneAfterStep(Lane& L, const FwdArgs& a, const DevSurface* geo);

// One iteration of the propagation loop (Propagator.ipp:25-210): step, target
// status, the actor's finalisation checks, aborters, next target.  Rev: the
// reverse filter, propagating backwards.
template <bool Mixed, bool Sympy = false, bool Rev = false>
__device__ inline int laneStep(Lane& L, const FwdArgs& a, const DevSurface* geo) {
  const DevOpts& o = a.o;
  Track& T = L.T;
  if (L.iter >= o.maxSteps) {
    L.status = ACTS_KF_STATUS_STEP_COUNT_LIMIT;
    return kPhaseFinished;
  }
  constexpr double propDir = Rev ? -1.0 : 1.0;
  if (!Rev && a.skipAfter > 0 && !L.noTransport && !L.replaying && !L.noSkip &&
      L.iter - L.lastSurfIter >= static_cast<uint32_t>(a.skipAfter)) {
    a.ckpt[static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x] = L;
    L.noTransport = true;
  }
  ++L.iter;
  const bool transport = !L.noTransport;
  int sst;
  if constexpr (Mixed && Sympy)
    sst = rkStepSympyMixed(T, a.F, o, propDir, transport);
  else if constexpr (Mixed)
    sst = rkStepMixed(T, a.F, o, propDir, transport);
  else if constexpr (Sympy)
    sst = rkStepSympy(T, a.F, o, propDir, transport);
  else
    sst = rkStep(T, a.F, o, propDir, transport);
  if (sst) {
    L.status = sst;
    return kPhaseFinished;
  }
  return laneAfterStep<Mixed, Sympy, Rev>(L, a, geo);
}

// The actor's finalisation checks (KalmanFitter.hpp:395-420): all
// measurements handled, the path limit, and for the reverse filter its
// target (each check updates the actor step size, in this order).
template <bool Rev, bool Sympy>
__device__ inline bool actorFinished(Lane& L, const FwdArgs& a) {
  bool fin = (L.nMeasStates == L.nUnique) || pathAbort(L.T, L.kfLimit, a.o.surfaceTolerance);
  if constexpr (Rev) {
    if (a.rv.hasTarget) fin = reverseTargetReached<Sympy>(L, a) || fin;
  }
  return fin;
}

// The rest of the propagation-loop iteration after an accepted step: target
// status, the actor's finalisation checks, aborters, next target.
template <bool Mixed, bool Sympy, bool Rev>
__device__ inline int laneAfterStep(Lane& L, const FwdArgs& a, const DevSurface* geo) {
  const DevOpts& o = a.o;
  Track& T = L.T;
  T.cNav = kNotSet;
  T.cActor = kNotSet;
  if (L.nv.tSurf >= 0) {
    const int ps = updateSurfaceStatus<Rev>(T, geo[L.nv.tSurf], L.nv.tIdx, o.surfaceTolerance,
                                            o.navigator != ACTS_KF_NAVIGATOR_DIRECT);
    if (ps == kOnSurface && !L.nv.navBreak) L.reached = L.nv.tSurf;
    if (ps != kReachable) L.nv.tSurf = -1;
  }
  if (L.reached >= 0) return kPhaseSurface;
  // act without a current surface: only the finalisation checks
  if (!L.finished) {
    if (L.kfLimit == DBL_MAX) loopProtection<Mixed>(T, a.F, o, L.kfLimit, true);
    L.finished = actorFinished<Rev, Sympy>(L, a);
  }
  if (L.finished || pathAbort(T, L.propLimit, o.surfaceTolerance)) return kPhaseFinished;
  if (L.nv.tSurf < 0) {
    T.cNav = kNotSet;
    L.status = getNextTarget<Rev>(L.nv, T, geo, o, L.seq, layerCtx(a, L.q), -1);
    if (L.status) return kPhaseFinished;
  }
  return kPhaseStepping;
}

// KalmanFitter::Actor::filter on the reached surface (KalmanFitter.hpp:
// 339-671), then the rest of the propagation-loop iteration.  The forward
// filter writes each state's record; the reverse filter (Rev) writes none
// and keeps only its last measurement update (RevArgs::rec).  Material is
// split by direction (ISurfaceMaterial.cpp:18-31, split factor 1): forwards
// after the update, backwards before it; on the reverse filter's start
// surface only post-update material counts, which backwards is none
// (PointwiseMaterialInteraction.cpp:16-29).
template <bool Mixed, bool Sympy = false, bool Rev = false>
__device__ inline int laneFilter(Lane& L, const FwdArgs& a, const DevSurface* geo,
                                 const DevSlab* slabs) {
  const DevBatch& in = a.in;
  const DevOpts& o = a.o;
  const DevScratch& sc = a.sc;
  Track& T = L.T;
  const int reached = L.reached;
  const int64_t q = L.q;
  L.reached = -1;
  L.lastSurfIter = L.iter;
  L.replaying = false;
  if (!L.finished) {
    if (L.kfLimit == DBL_MAX) loopProtection<Mixed>(T, a.F, o, L.kfLimit, true);
    const DevSurface& s = geo[reached];
    // the measurement on this surface: the first one in input order; when
    // all of the track's measurement surfaces differ and the measurements
    // come in the order of the surface sequence, it is the one after the
    // last match
    int mi = -1;
    if (!Rev && L.nUnique == L.nMeas && L.mNext < L.nMeas &&
        in.measSurf[L.m0 + L.mNext] == reached) {
      mi = L.mNext;
    } else {
      for (int i = 0; i < L.nMeas; ++i) {
        const int sm = in.measSurf[L.m0 + i];
        if (mi < 0 && sm == reached) mi = i;
      }
    }
    if (mi >= 0) L.mNext = mi + 1;
    const bool hasMat = s.mat >= 0;
    const bool matHere = hasMat && !(Rev && reached == L.startSurf);
    const bool sensitive = (s.flags & ACTS_KF_SURFACE_SENSITIVE) != 0;
    const bool preceding = L.nMeasStates > 0;
    if (mi >= 0 || (preceding && sensitive) || hasMat) {
      double l0 = 0, l1 = 0;
      const bool onS = globalToLocal(s, T.pos, l0, l1);
      double J[36];
      transportToBound<Sympy>(T, s, onS, J);
      if (!onS) {
        L.status = ACTS_KF_STATUS_NOT_ON_SURFACE;
      } else if (!Rev && L.nStates >= L.nv.seqLen) {
        // only the layer navigator can find more states than slots
        L.status = ACTS_KF_STATUS_STATE_CAPACITY;
      } else if (!Rev && L.nStates >= sc.maxStates) {
        L.status = ACTS_KF_STATUS_PROPAGATION_FAILURE;
      } else {
        // going backwards, a measurement surface's material comes before
        // its update (the predicted state includes it)
        if (Rev && mi >= 0 && matHere)
          L.status = materialInteraction(T, s, slabs[s.mat], a.pc, o, -1.0);
      }
      if (!L.status) {
        // bound parameters (transformFreeToBoundParameters)
        double x[6];
        x[0] = l0;
        x[1] = l1;
        x[2] = atan2(T.dir.y, T.dir.x);
        x[3] = atan2(perp(T.dir), T.dir.z);
        x[4] = T.qop;
        x[5] = T.time;
        const int k = L.nStates;
        uint32_t flags = ACTS_KF_STATE_HAS_PARAMETERS;
        if (hasMat) flags |= ACTS_KF_STATE_HAS_MATERIAL;
        if constexpr (!Rev) {
#pragma unroll
          for (int i = 0; i < 6; ++i) sd(sc, k, kPred + i, q) = x[i];
#pragma unroll
          for (int i = 0; i < 21; ++i) sd(sc, k, kPCov + i, q) = T.cov[i];
#pragma unroll
          for (int i = 0; i < 36; ++i) sd(sc, k, kJac + i, q) = J[i];
          sd(sc, k, kPath, q) = T.pathAcc;
        }
        float chi2f = 0.f;
        bool measured = false;
        if (mi >= 0) {
          flags |= ACTS_KF_STATE_HAS_MEASUREMENT;
          MeasDev M;
          loadMeas(in, L.m0 + mi, M);  // its subspace was checked by laneInit
          bool outlier = false;
          if (o.chi2Cut > 0) outlier = predictedChi2Meas(x, T.cov, M) > o.chi2Cut;
          if (!outlier) {
            double f[6], C[21], chi2 = 0;
            const int us = gainUpdateMeas(x, T.cov, M, o.useJoseph != 0, f, C, chi2);
            if (us) {
              L.status = us;
            } else {
              measured = true;
              chi2f = static_cast<float>(chi2);
              if constexpr (Rev) {
                // the reverse filter's latest measurement state
                double* r = a.rv.rec + kRevRecord * q;
#pragma unroll
                for (int i = 0; i < 6; ++i) r[i] = f[i];
#pragma unroll
                for (int i = 0; i < 21; ++i) r[6 + i] = C[i];
                r[27] = static_cast<double>(reached);
              } else {
#pragma unroll
                for (int i = 0; i < 6; ++i) sd(sc, k, kFilt + i, q) = f[i];
#pragma unroll
                for (int i = 0; i < 21; ++i) sd(sc, k, kFCov + i, q) = C[i];
              }
              // EigenStepper::update with the filtered parameters
              const double sinTh = sin(f[3]);
              T.dir = mk(cos(f[2]) * sinTh, sin(f[2]) * sinTh, cos(f[3]));
              T.pos = localToGlobal(s, f[0], f[1]);
              T.time = f[5];
              T.qop = f[4];
#pragma unroll
              for (int i = 0; i < 21; ++i) T.cov[i] = C[i];
              double b2f[6];
              positionJacobians(s, T.pos, b2f, nullptr);
#pragma unroll
              for (int i = 0; i < 6; ++i) T.jp[i] = b2f[i];
              directionB2F(T.dir, T.jd);
              if constexpr (Sympy) sympyResetM(T);
              ++L.nMeasStates;
              if (L.firstMeas < 0) L.firstMeas = k;
              L.lastMeas = k;
            }
          } else {
            flags |= ACTS_KF_STATE_IS_OUTLIER;
          }
        } else if (sensitive && preceding) {
          flags |= ACTS_KF_STATE_IS_HOLE;
        }
        if (!L.status) {
          if constexpr (!Rev) {
            if (!measured) {
#pragma unroll
              for (int i = 0; i < 6; ++i) sd(sc, k, kFilt + i, q) = x[i];
#pragma unroll
              for (int i = 0; i < 21; ++i) sd(sc, k, kFCov + i, q) = T.cov[i];
            }
            sd(sc, k, kChi2, q) = chi2f;
            si(sc, k, kISurf, q) = reached;
            si(sc, k, kIFlags, q) = static_cast<int32_t>(flags);
            si(sc, k, kIMeas, q) = mi;
          }
          L.nStates = k + 1;
          if (mi >= 0) L.tip = k;
          // after the update going forwards; a material-only or hole state
          // takes the full material in either direction
          if (matHere && (!Rev || mi < 0))
            L.status = materialInteraction(T, s, slabs[s.mat], a.pc, o, Rev ? -1.0 : 1.0);
        }
      }
    }
    if (L.status) return kPhaseFinished;
    L.finished = actorFinished<Rev, Sympy>(L, a);
  }
  if (L.finished || pathAbort(T, L.propLimit, o.surfaceTolerance)) return kPhaseFinished;
  T.cNav = kNotSet;
  L.status = getNextTarget<Rev>(L.nv, T, geo, o, L.seq, layerCtx(a, q), reached);
  return L.status ? kPhaseFinished : kPhaseStepping;
}

// End of the forward pass of a track: counters and the smoother's inputs;
// of the reverse pass: reverseFinish.
template <bool Rev = false>
__device__ inline void laneFinish(Lane& L, const FwdArgs& a) {
  if constexpr (Rev) {
    reverseFinish(L, a);
    return;
  }
  const int64_t t = L.t;
  int status = L.status;
  if (a.out.nSteps) a.out.nSteps[t] = L.T.nSteps;
  if (a.out.nTrials) a.out.nTrials[t] = L.T.nTrials;
  if (!status && L.nMeasStates == 0) status = ACTS_KF_STATUS_NO_MEASUREMENT_FOUND;
  const int64_t q = L.q;
  a.sc.track[q] = status;
  a.sc.track[a.sc.stride + q] = L.tip;
  a.sc.track[2 * a.sc.stride + q] = L.firstMeas;
  a.sc.track[3 * a.sc.stride + q] = L.lastMeas;
}

// An output slot that holds no state: surface and measurement -1, flags 0,
// every number 0.
__device__ inline void clearSlot(const DevOut& out, int64_t slot) {
  if (out.smoothed)
    for (int i = 0; i < 6; ++i) out.smoothed[6 * slot + i] = 0.0;
  if (out.smoothedCov)
    for (int i = 0; i < 36; ++i) out.smoothedCov[36 * slot + i] = 0.0;
  if (out.filtered)
    for (int i = 0; i < 6; ++i) out.filtered[6 * slot + i] = 0.0;
  if (out.filteredCov)
    for (int i = 0; i < 36; ++i) out.filteredCov[36 * slot + i] = 0.0;
  if (out.predicted)
    for (int i = 0; i < 6; ++i) out.predicted[6 * slot + i] = 0.0;
  if (out.predictedCov)
    for (int i = 0; i < 36; ++i) out.predictedCov[36 * slot + i] = 0.0;
  if (out.jacobian)
    for (int i = 0; i < 36; ++i) out.jacobian[36 * slot + i] = 0.0;
  if (out.stateSurface) out.stateSurface[slot] = -1;
  if (out.stateFlags) out.stateFlags[slot] = 0u;
  if (out.statePath) out.statePath[slot] = 0.0;
  if (out.stateChi2) out.stateChi2[slot] = 0.f;
  if (out.stateMeas) out.stateMeas[slot] = -1;
}

// Expand a packed symmetric matrix into a row-major 6x6 output.
__device__ inline void storeSym(double* dst, const double* p) {
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) dst[6 * i + j] = p[sym(i, j)];
}

// GainMatrixSmoother::operator() from the tip backwards, the track summary
// (calculateTrackQuantities) and the per-state outputs.
// `valid` is false for the padding lanes of the last wavefront: they take
// part in the wavefront-wide reduction but touch no memory.
// MbfSmoother::calculateSmoothed (MbfSmoother.cpp:16-27) with the carried
// Lambda (full 6x6) and lambda: Cs = Cf - (Cf Lambda) Cf, xs = xf - Cf lambda
__device__ inline void mbfSmoothed(const double xf[6], const double Cf[21], const double L[36],
                                   const double l[6], double xs[6], double Cs[36]) {
  double CfL[36];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double acc = Cf[sym(i, 0)] * L[j];
#pragma unroll
      for (int m = 1; m < 6; ++m) acc = acc + Cf[sym(i, m)] * L[6 * m + j];
      CfL[6 * i + j] = acc;
    }
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double acc = CfL[6 * i] * Cf[sym(0, j)];
#pragma unroll
      for (int m = 1; m < 6; ++m) acc = acc + CfL[6 * i + m] * Cf[sym(m, j)];
      Cs[6 * i + j] = Cf[sym(i, j)] - acc;
    }
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double acc = Cf[sym(i, 0)] * l[0];
#pragma unroll
    for (int m = 1; m < 6; ++m) acc = acc + Cf[sym(i, m)] * l[m];
    xs[i] = xf[i] - acc;
  }
  normalizePhiTheta(xs[2], xs[3]);
}

// (F^T A) F for the full 6x6 A, in place
__device__ inline void mbfTransport(const double F[36], double A[36]) {
  double T[36];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double acc = F[i] * A[j];
#pragma unroll
      for (int m = 1; m < 6; ++m) acc = acc + F[6 * m + i] * A[6 * m + j];
      T[6 * i + j] = acc;
    }
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double acc = T[6 * i] * F[j];
#pragma unroll
      for (int m = 1; m < 6; ++m) acc = acc + T[6 * i + m] * F[6 * m + j];
      A[6 * i + j] = acc;
    }
}

// The Lambda / lambda update of state k on the way to state k-1
// (MbfSmoother.cpp:29-48, detail/MbfSmootherImpl.hpp:17-57): through the
// measurement update if state k is a measurement, then through its
// transport jacobian F.
__device__ inline void mbfUpdate(int64_t q, int64_t t, int k, uint32_t flags, const DevBatch& in,
                                 const DevScratch& sc, double L[36], double l[6]) {
  double F[36];
#pragma unroll
  for (int i = 0; i < 36; ++i) F[i] = sd(sc, k, kJac + i, q);
  const bool isMeasurement =
      (flags & ACTS_KF_STATE_HAS_MEASUREMENT) && !(flags & ACTS_KF_STATE_IS_OUTLIER);
  MeasDev M;
  if (isMeasurement) loadMeas(in, in.measOff[t] + si(sc, k, kIMeas, q), M);
  if (isMeasurement && !isLoc01(M)) {
    // a measurement of another subspace: the same recursion with the
    // projector's indices a (, b) in place of loc0 (, loc1)
    double P[21], xp[6];
#pragma unroll
    for (int i = 0; i < 21; ++i) P[i] = sd(sc, k, kPCov + i, q);
#pragma unroll
    for (int i = 0; i < 6; ++i) xp[i] = sd(sc, k, kPred + i, q);
    const bool two = M.n == 2;
    double Paa = 0, Pab = 0, Pba = 0, Pbb = 0, xa = 0, xb = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      xa = (M.a == i) ? xp[i] : xa;
      xb = (M.b == i) ? xp[i] : xb;
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const double pij = P[sym(i, j)];
        Paa = (M.a == i && M.a == j) ? pij : Paa;
        Pab = (M.a == i && M.b == j) ? pij : Pab;
        Pba = (M.b == i && M.a == j) ? pij : Pba;
        Pbb = (M.b == i && M.b == j) ? pij : Pbb;
      }
    }
    double Si[2][2] = {{0, 0}, {0, 0}};
    if (!two) {
      Si[0][0] = 1.0 / (Paa + M.V[0]);
    } else {
      double o[4];
      inverse2(Paa + M.V[0], Pab + M.V[1], Pba + M.V[2], Pbb + M.V[3], o);
      Si[0][0] = o[0];
      Si[0][1] = o[1];
      Si[1][0] = o[2];
      Si[1][1] = o[3];
    }
    // K = (P H^T) S^-1, CHat = I - K H
    double CH[36];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      double pia = 0, pib = 0;
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        pia = (M.a == j) ? P[sym(i, j)] : pia;
        pib = (M.b == j) ? P[sym(i, j)] : pib;
      }
      double k0 = pia * Si[0][0], k1 = pia * Si[0][1];
      if (two) {
        k0 = k0 + pib * Si[1][0];
        k1 = k1 + pib * Si[1][1];
      }
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const double id = (i == j) ? 1.0 : 0.0;
        CH[6 * i + j] = (M.a == j) ? id - k0 : ((two && M.b == j) ? id - k1 : id);
      }
    }
    const double y0 = M.v[0] - xa, y1 = two ? M.v[1] - xb : 0.0;
    // Lambda~ = (H^T S^-1) H + (CHat^T Lambda) CHat
    mbfTransport(CH, L);
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const int r = (M.a == i) ? 0 : ((two && M.b == i) ? 1 : -1);
        const int c = (M.a == j) ? 0 : ((two && M.b == j) ? 1 : -1);
        if (r >= 0 && c >= 0) {
          const double sij = (r == 0) ? (c == 0 ? Si[0][0] : Si[0][1])
                                      : (c == 0 ? Si[1][0] : Si[1][1]);
          L[6 * i + j] = sij + L[6 * i + j];
        }
      }
    // lambda~ = (-H^T S^-1) y + CHat^T lambda
    double lt[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      double acc = CH[i] * l[0];
#pragma unroll
      for (int m = 1; m < 6; ++m) acc = acc + CH[6 * m + i] * l[m];
      lt[i] = acc;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const int r = (M.a == i) ? 0 : ((two && M.b == i) ? 1 : -1);
      if (r >= 0) {
        const double s0 = (r == 0) ? Si[0][0] : Si[1][0];
        const double s1 = (r == 0) ? Si[0][1] : Si[1][1];
        double acc = -s0 * y0;
        if (two) acc = acc + -s1 * y1;
        lt[i] = acc + lt[i];
      }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) l[i] = lt[i];
  } else if (isMeasurement) {
    double P[21];
#pragma unroll
    for (int i = 0; i < 21; ++i) P[i] = sd(sc, k, kPCov + i, q);
    const int64_t mg = in.measOff[t] + si(sc, k, kIMeas, q);
    const double m0 = in.measVal[2 * mg], m1 = in.measVal[2 * mg + 1];
    const double* V = in.measCov + 4 * mg;
    // S = H P H^T + V and Eigen's closed-form 2x2 inverse
    const double S00 = P[sym(0, 0)] + V[0], S01 = P[sym(0, 1)] + V[1];
    const double S10 = P[sym(1, 0)] + V[2], S11 = P[sym(1, 1)] + V[3];
    const double det = S00 * S11 - S10 * S01;
    const double invdet = 1.0 / det;
    const double Si[2][2] = {{S11 * invdet, -S01 * invdet}, {-S10 * invdet, S00 * invdet}};
    // CHat = I - (P H^T S^-1) H
    double CH[36];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const double id = (i == j) ? 1.0 : 0.0;
        CH[6 * i + j] =
            j < 2 ? id - (P[sym(i, 0)] * Si[0][j] + P[sym(i, 1)] * Si[1][j]) : id;
      }
    const double y0 = m0 - sd(sc, k, kPred + 0, q), y1 = m1 - sd(sc, k, kPred + 1, q);
    // Lambda~ = (H^T S^-1) H + (CHat^T Lambda) CHat, by the transport helper
    // with CHat in place of F
    mbfTransport(CH, L);
    L[0] = Si[0][0] + L[0];
    L[1] = Si[0][1] + L[1];
    L[6] = Si[1][0] + L[6];
    L[7] = Si[1][1] + L[7];
    // lambda~ = (-H^T S^-1) y + CHat^T lambda
    double lt[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      double acc = CH[i] * l[0];
#pragma unroll
      for (int m = 1; m < 6; ++m) acc = acc + CH[6 * m + i] * l[m];
      lt[i] = acc;
    }
    lt[0] = (-Si[0][0] * y0 + -Si[0][1] * y1) + lt[0];
    lt[1] = (-Si[1][0] * y0 + -Si[1][1] * y1) + lt[1];
#pragma unroll
    for (int i = 0; i < 6; ++i) l[i] = lt[i];
  }
  mbfTransport(F, L);
  double ln[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double acc = F[i] * l[0];
#pragma unroll
    for (int m = 1; m < 6; ++m) acc = acc + F[6 * m + i] * l[m];
    ln[i] = acc;
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) l[i] = ln[i];
}

// Mbf = false: GainMatrixSmoother (RTS); true: MbfSmoother.  `rev`: the
// track is filtered in reverse instead of smoothed (KalmanFitter.hpp:897-
// 945); its smoothed slots are NaN here, the reverse pass fills the first
// and last measurement states.
template <bool Mbf>
__device__ inline void smoothTrack(const int64_t q, const int64_t t, const bool valid,
                                   const DevBatch& in,
                                   const DevOut& out, const DevScratch& sc, const bool rev) {
  int status = valid ? sc.track[q] : 0;
  const int tip = (!valid || status) ? -1 : sc.track[sc.stride + q];
  // with a reference surface: keep the smoothed first and last measurement
  // states for the extrapolation
  const int firstMeas = (valid && sc.ref) ? sc.track[2 * sc.stride + q] : -2;
  const int lastMeas = (valid && sc.ref) ? sc.track[3 * sc.stride + q] : -2;
  int nHoles = 0, nOut = 0, nMeasT = 0, ndf = 0;
  float chi2sum = 0.f;
  // All lanes walk the same state index k (from the wavefront's largest tip
  // down), which keeps the lanes of a wavefront converged.
  int kTop = tip;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) kTop = max(kTop, __shfl_xor(kTop, off));
  const int64_t base = valid ? in.seqOff[t] : 0;
  double xsN[6], CsN[21];
  double L[36], l[6];  // MbfSmoother's Lambda-hat and lambda-hat
  if constexpr (Mbf) {
#pragma unroll
    for (int i = 0; i < 36; ++i) L[i] = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) l[i] = 0.0;
  }
  bool active = tip >= 0;
  for (int k = kTop; k >= 0; --k) {
    if (!active || k > tip) continue;
    double xf[6], Cf[21], xs[6], Cs[21], CsFull[36];
#pragma unroll
    for (int i = 0; i < 6; ++i) xf[i] = sd(sc, k, kFilt + i, q);
#pragma unroll
    for (int i = 0; i < 21; ++i) Cf[i] = sd(sc, k, kFCov + i, q);
    if (rev) {
      const double nan = __builtin_nan("");
#pragma unroll
      for (int i = 0; i < 6; ++i) xs[i] = nan;
#pragma unroll
      for (int i = 0; i < 21; ++i) Cs[i] = nan;
#pragma unroll
      for (int i = 0; i < 36; ++i) CsFull[i] = nan;
    } else if constexpr (Mbf) {
      mbfSmoothed(xf, Cf, L, l, xs, CsFull);
#pragma unroll
      for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = i; j < 6; ++j) Cs[sym(i, j)] = CsFull[6 * i + j];
    } else if (k == tip) {
#pragma unroll
      for (int i = 0; i < 6; ++i) xs[i] = xf[i];
#pragma unroll
      for (int i = 0; i < 21; ++i) Cs[i] = Cf[i];
    } else {
      double xpN[6], CpN[21], JN[36];
#pragma unroll
      for (int i = 0; i < 6; ++i) xpN[i] = sd(sc, k + 1, kPred + i, q);
#pragma unroll
      for (int i = 0; i < 21; ++i) CpN[i] = sd(sc, k + 1, kPCov + i, q);
#pragma unroll
      for (int i = 0; i < 36; ++i) JN[i] = sd(sc, k + 1, kJac + i, q);
      const int r = smoothStep(xf, Cf, xpN, CpN, xsN, CsN, JN, xs, Cs);
      if (r) {
        status = r;
        active = false;
        continue;
      }
    }
    if (!rev && (k == firstMeas || k == lastMeas)) {
      double* r = sc.ref + kRefComps * q + (k == firstMeas ? 0 : 27);
#pragma unroll
      for (int i = 0; i < 6; ++i) r[i] = xs[i];
#pragma unroll
      for (int i = 0; i < 21; ++i) r[6 + i] = Cs[i];
      if (k == firstMeas && k == lastMeas) {
#pragma unroll
        for (int i = 0; i < 27; ++i) r[27 + i] = r[i];
      }
    }
    const uint32_t flags = static_cast<uint32_t>(si(sc, k, kIFlags, q));
    const float c2 = static_cast<float>(sd(sc, k, kChi2, q));
    if (flags & ACTS_KF_STATE_IS_HOLE) {
      ++nHoles;
    } else if (flags & ACTS_KF_STATE_IS_OUTLIER) {
      ++nOut;
    } else if (flags & ACTS_KF_STATE_HAS_MEASUREMENT) {
      ++nMeasT;
      chi2sum += c2;
      ndf += in.measDims ? in.measDims[in.measOff[t] + si(sc, k, kIMeas, q)] : 2;
    }
    const int64_t slot = base + k;
    if (out.smoothed)
      for (int i = 0; i < 6; ++i) out.smoothed[6 * slot + i] = xs[i];
    if (out.smoothedCov) {
      if constexpr (Mbf) {
        for (int i = 0; i < 36; ++i) out.smoothedCov[36 * slot + i] = CsFull[i];
      } else {
        storeSym(out.smoothedCov + 36 * slot, Cs);
      }
    }
    if (out.filtered)
      for (int i = 0; i < 6; ++i) out.filtered[6 * slot + i] = xf[i];
    if (out.filteredCov) storeSym(out.filteredCov + 36 * slot, Cf);
    if (out.predicted)
      for (int i = 0; i < 6; ++i) out.predicted[6 * slot + i] = sd(sc, k, kPred + i, q);
    if (out.predictedCov) {
      double P[21];
      for (int i = 0; i < 21; ++i) P[i] = sd(sc, k, kPCov + i, q);
      storeSym(out.predictedCov + 36 * slot, P);
    }
    if (out.jacobian)
      for (int i = 0; i < 36; ++i) out.jacobian[36 * slot + i] = sd(sc, k, kJac + i, q);
    if (out.stateSurface) out.stateSurface[slot] = si(sc, k, kISurf, q);
    if (out.stateFlags) out.stateFlags[slot] = flags;
    if (out.statePath) out.statePath[slot] = sd(sc, k, kPath, q);
    if (out.stateChi2) out.stateChi2[slot] = c2;
    if (out.stateMeas) {
      const int32_t mi = si(sc, k, kIMeas, q);
      out.stateMeas[slot] = mi >= 0 ? static_cast<int32_t>(in.measOff[t] + mi) : -1;
    }
    if constexpr (Mbf) {
      if (k > 0 && !rev) mbfUpdate(q, t, k, flags, in, sc, L, l);
    } else {
#pragma unroll
      for (int i = 0; i < 6; ++i) xsN[i] = xs[i];
#pragma unroll
      for (int i = 0; i < 21; ++i) CsN[i] = Cs[i];
    }
  }
  if (!valid) return;
  // slots past the track's states (all of them for a failed track) get the
  // values of an empty slot, so that every output slot is defined
  const int seqLen = static_cast<int>(in.seqOff[t + 1] - base);
  for (int k = status ? 0 : tip + 1; k < seqLen; ++k) clearSlot(out, base + k);
  // the status after smoothing, for the extrapolation kernel
  sc.track[q] = status;
  if (out.status) out.status[t] = status;
  if (out.nStates) out.nStates[t] = status ? 0 : tip + 1;
  if (out.chi2) out.chi2[t] = status ? 0.f : chi2sum;
  if (out.ndf) out.ndf[t] = status ? 0 : ndf;
  if (out.nMeas) out.nMeas[t] = status ? 0 : nMeasT;
  if (out.nHoles) out.nHoles[t] = status ? 0 : nHoles;
  if (out.nOutliers) out.nOutliers[t] = status ? 0 : nOut;
  if (out.reversed) out.reversed[t] = 0;
}

}  // namespace kf
