// kf_kernel.hpp -- the fit kernel: one lane per track.
//
// Forward pass: Propagator::propagate with the KalmanFitter actor
// (Propagator.ipp:25-210, KalmanFitter.hpp:339-671), restructured so that a
// lane that reaches its next surface waits for the other lanes of its
// wavefront before the (expensive) surface update runs for all of them
// together.  The order of operations seen by each track is unchanged.
// Backward pass (a separate kernel, so that each pass gets its own register
// allocation): GainMatrixSmoother from the tip (GainMatrixSmoother.hpp:43-99)
// and calculateTrackQuantities (TrackHelpers.hpp:421-445).
#pragma once

#include "kf_fit.hpp"

namespace kf {

// Scratch of lane g: wavefront tile w = g / 64 holds its states contiguously,
// element c of state k at ((w * maxStates + k) * comps + c) * 64 + g % 64.
// A wavefront's stores of one component are 512 consecutive bytes, and the
// states of one wavefront stay within a few pages.
__device__ inline double& sd(const DevScratch& sc, int k, int c, int64_t g) {
  const int64_t w = g >> 6;
  return sc.d[((w * sc.maxStates + k) * kScratchComps + c) * 64 + (g & 63)];
}
__device__ inline int32_t& si(const DevScratch& sc, int k, int c, int64_t g) {
  const int64_t w = g >> 6;
  return sc.i[((w * sc.maxStates + k) * kIScratchComps + c) * 64 + (g & 63)];
}

// t: the track (input/output index); g: the lane's scratch column.
__device__ inline void forwardTrack(const int64_t t, const int64_t g, const DevBatch& in,
                                   const DevOut& out, const DevScratch& sc,
                                   const DevSurface* __restrict__ geo,
                                   const DevSlab* __restrict__ slabs,
                                   const DevField& F, const DevOpts& o,
                                   const PhysConst& pc, const int32_t nSurfaces) {
  const int64_t s0 = in.seqOff[t];
  const int seqLen = static_cast<int>(in.seqOff[t + 1] - s0);
  const int64_t m0 = in.measOff[t];
  const int nMeas = static_cast<int>(in.measOff[t + 1] - m0);
  auto seqAt = [&](int i) { return in.seqSurf[s0 + i]; };
  // surface indices come from the caller: reject the track if any is out of
  // range instead of reading outside the geometry table
  bool badIndex = false;
  for (int i = 0; i < seqLen; ++i) {
    const int sid = seqAt(i);
    badIndex = badIndex || sid < 0 || sid >= nSurfaces;
  }
  if (badIndex) {
    sc.track[g] = ACTS_KF_STATUS_PROPAGATION_FAILURE;
    sc.track[sc.stride + g] = -1;
    if (out.nSteps) out.nSteps[t] = 0;
    if (out.nTrials) out.nTrials[t] = 0;
    return;
  }

  // unique measurement surfaces (KalmanFitter.hpp:750-755, first one wins)
  int nUnique = 0;
  for (int i = 0; i < nMeas; ++i) {
    const int si_ = in.measSurf[m0 + i];
    bool dup = false;
    for (int j = 0; j < i; ++j) dup = dup || (in.measSurf[m0 + j] == si_);
    nUnique += dup ? 0 : 1;
  }

  // EigenStepper::initialize (EigenStepper.ipp:37-70)
  Track T;
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
    T.pathAcc = 0;
    T.nSteps = 0;
    T.nTrials = 0;
    T.cNav = kNotSet;
    T.cActor = kNotSet;
    T.cUser = o.maxStepSize;
    T.cAcc = o.initialStepSize;
  }

  Nav nv;
  nv.seqPos = -1;
  nv.seqLen = seqLen;
  nv.navBreak = false;
  nv.tSurf = -1;
  nv.tIdx = 0;

  double propLimit = o.pathLimit;
  loopProtection(T, F, o, propLimit, false);
  double kfLimit = DBL_MAX;
  int nMeasStates = 0, nStates = 0, tip = -1;
  bool finished = false;
  int status = 0;

  // pre-propagation act (the start surface carries neither material nor a
  // measurement, so only the finalisation checks run)
  if (kfLimit == DBL_MAX) loopProtection(T, F, o, kfLimit, true);
  {
    const bool complete = nMeasStates == nUnique;
    const bool pathReached = pathAbort(T, kfLimit, o.surfaceTolerance);
    finished = complete || pathReached;
  }
  bool done = finished || pathAbort(T, propLimit, o.surfaceTolerance);
  if (!done) {
    status = getNextTarget(nv, T, geo, o, seqAt);
    done = status != 0;
  }
  uint32_t iter = 0;

  while (!done) {
    // ---- stepping until this lane reaches a surface or stops ----
    int reached = -1;
    while (true) {
      if (iter >= o.maxSteps) {
        status = ACTS_KF_STATUS_STEP_COUNT_LIMIT;
        done = true;
        break;
      }
      ++iter;
      const int sst = rkStep(T, F, o);
      if (sst) {
        status = sst;
        done = true;
        break;
      }
      T.cNav = kNotSet;
      T.cActor = kNotSet;
      if (nv.tSurf >= 0) {
        const int ps = updateSurfaceStatus(T, geo[nv.tSurf], nv.tIdx, o.surfaceTolerance);
        if (ps == kOnSurface && !nv.navBreak) reached = nv.tSurf;
        if (ps != kReachable) nv.tSurf = -1;
      }
      if (reached >= 0) break;
      // act without a current surface: only the finalisation checks
      if (!finished) {
        if (kfLimit == DBL_MAX) loopProtection(T, F, o, kfLimit, true);
        const bool complete = nMeasStates == nUnique;
        const bool pathReached = pathAbort(T, kfLimit, o.surfaceTolerance);
        finished = complete || pathReached;
      }
      if (finished || pathAbort(T, propLimit, o.surfaceTolerance)) {
        done = true;
        break;
      }
      if (nv.tSurf < 0) {
        T.cNav = kNotSet;
        status = getNextTarget(nv, T, geo, o, seqAt);
        if (status) {
          done = true;
          break;
        }
      }
    }
    if (reached < 0) break;

    // ---- KalmanFitter::Actor::filter on the reached surface ----
    if (!finished) {
      if (kfLimit == DBL_MAX) loopProtection(T, F, o, kfLimit, true);
      const DevSurface& s = geo[reached];
      int mi = -1;
      for (int i = 0; i < nMeas; ++i)
        if (mi < 0 && in.measSurf[m0 + i] == reached) mi = i;
      const bool hasMat = s.mat >= 0;
      const bool sensitive = (s.flags & ACTS_KF_SURFACE_SENSITIVE) != 0;
      const bool preceding = nMeasStates > 0;
      if (mi >= 0 || (preceding && sensitive) || hasMat) {
        double l0 = 0, l1 = 0;
        const bool onS = globalToLocal(s, T.pos, l0, l1);
        double J[36];
        transportToBound(T, s, onS, J);
        if (!onS) {
          status = ACTS_KF_STATUS_NOT_ON_SURFACE;
        } else if (nStates >= sc.maxStates) {
          status = ACTS_KF_STATUS_PROPAGATION_FAILURE;
        } else {
          // bound parameters (transformFreeToBoundParameters)
          double x[6];
          x[0] = l0;
          x[1] = l1;
          x[2] = atan2(T.dir.y, T.dir.x);
          x[3] = atan2(perp(T.dir), T.dir.z);
          x[4] = T.qop;
          x[5] = T.time;
          const int k = nStates;
          uint32_t flags = ACTS_KF_STATE_HAS_PARAMETERS;
          if (hasMat) flags |= ACTS_KF_STATE_HAS_MATERIAL;
#pragma unroll
          for (int i = 0; i < 6; ++i) sd(sc, k, kPred + i, g) = x[i];
#pragma unroll
          for (int i = 0; i < 21; ++i) sd(sc, k, kPCov + i, g) = T.cov[i];
#pragma unroll
          for (int i = 0; i < 36; ++i) sd(sc, k, kJac + i, g) = J[i];
          sd(sc, k, kPath, g) = T.pathAcc;
          float chi2f = 0.f;
          bool measured = false;
          if (mi >= 0) {
            flags |= ACTS_KF_STATE_HAS_MEASUREMENT;
            const double* mv = in.measVal + 2 * (m0 + mi);
            const double* mc = in.measCov + 4 * (m0 + mi);
            const double mvv[2] = {mv[0], mv[1]};
            const double mcc[4] = {mc[0], mc[1], mc[2], mc[3]};
            bool outlier = false;
            if (o.chi2Cut > 0) outlier = predictedChi2(x, T.cov, mvv, mcc) > o.chi2Cut;
            if (!outlier) {
              double f[6], C[21], chi2 = 0;
              const int us = gainUpdate(x, T.cov, mvv, mcc, o.useJoseph != 0, f, C, chi2);
              if (us) {
                status = us;
              } else {
                measured = true;
                chi2f = static_cast<float>(chi2);
#pragma unroll
                for (int i = 0; i < 6; ++i) sd(sc, k, kFilt + i, g) = f[i];
#pragma unroll
                for (int i = 0; i < 21; ++i) sd(sc, k, kFCov + i, g) = C[i];
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
                ++nMeasStates;
              }
            } else {
              flags |= ACTS_KF_STATE_IS_OUTLIER;
            }
          } else if (sensitive && preceding) {
            flags |= ACTS_KF_STATE_IS_HOLE;
          }
          if (!status) {
            if (!measured) {
#pragma unroll
              for (int i = 0; i < 6; ++i) sd(sc, k, kFilt + i, g) = x[i];
#pragma unroll
              for (int i = 0; i < 21; ++i) sd(sc, k, kFCov + i, g) = T.cov[i];
            }
            sd(sc, k, kChi2, g) = chi2f;
            si(sc, k, kISurf, g) = reached;
            si(sc, k, kIFlags, g) = static_cast<int32_t>(flags);
            nStates = k + 1;
            if (mi >= 0) tip = k;
            if (hasMat) status = materialInteraction(T, s, slabs[s.mat], pc, o);
          }
        }
      }
      if (status) {
        done = true;
      } else {
        const bool complete = nMeasStates == nUnique;
        const bool pathReached = pathAbort(T, kfLimit, o.surfaceTolerance);
        finished = complete || pathReached;
      }
    }
    if (done) break;
    if (finished || pathAbort(T, propLimit, o.surfaceTolerance)) break;
    T.cNav = kNotSet;
    status = getNextTarget(nv, T, geo, o, seqAt);
    if (status) break;
  }

  if (out.nSteps) out.nSteps[t] = T.nSteps;
  if (out.nTrials) out.nTrials[t] = T.nTrials;
  if (!status && nMeasStates == 0) status = ACTS_KF_STATUS_NO_MEASUREMENT_FOUND;
  sc.track[g] = status;
  sc.track[sc.stride + g] = tip;
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
__device__ inline void smoothTrack(const int64_t t, const int64_t g, const bool valid,
                                   const DevBatch& in, const DevOut& out,
                                   const DevScratch& sc) {
  int status = valid ? sc.track[g] : 0;
  const int tip = (!valid || status) ? -1 : sc.track[sc.stride + g];
  int nHoles = 0, nOut = 0, nMeasT = 0, ndf = 0;
  float chi2sum = 0.f;
  // All lanes walk the same state index k (from the wavefront's largest tip
  // down), so every scratch load of a wavefront hits one contiguous row.
  int kTop = tip;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) kTop = max(kTop, __shfl_xor(kTop, off));
  const int64_t base = valid ? in.seqOff[t] : 0;
  double xsN[6], CsN[21];
  bool active = tip >= 0;
  for (int k = kTop; k >= 0; --k) {
    if (!active || k > tip) continue;
    double xf[6], Cf[21], xs[6], Cs[21];
#pragma unroll
    for (int i = 0; i < 6; ++i) xf[i] = sd(sc, k, kFilt + i, g);
#pragma unroll
    for (int i = 0; i < 21; ++i) Cf[i] = sd(sc, k, kFCov + i, g);
    if (k == tip) {
#pragma unroll
      for (int i = 0; i < 6; ++i) xs[i] = xf[i];
#pragma unroll
      for (int i = 0; i < 21; ++i) Cs[i] = Cf[i];
    } else {
      double xpN[6], CpN[21], JN[36];
#pragma unroll
      for (int i = 0; i < 6; ++i) xpN[i] = sd(sc, k + 1, kPred + i, g);
#pragma unroll
      for (int i = 0; i < 21; ++i) CpN[i] = sd(sc, k + 1, kPCov + i, g);
#pragma unroll
      for (int i = 0; i < 36; ++i) JN[i] = sd(sc, k + 1, kJac + i, g);
      const int r = smoothStep(xf, Cf, xpN, CpN, xsN, CsN, JN, xs, Cs);
      if (r) {
        status = r;
        active = false;
        continue;
      }
    }
    const uint32_t flags = static_cast<uint32_t>(si(sc, k, kIFlags, g));
    const float c2 = static_cast<float>(sd(sc, k, kChi2, g));
    if (flags & ACTS_KF_STATE_IS_HOLE) {
      ++nHoles;
    } else if (flags & ACTS_KF_STATE_IS_OUTLIER) {
      ++nOut;
    } else if (flags & ACTS_KF_STATE_HAS_MEASUREMENT) {
      ++nMeasT;
      chi2sum += c2;
      ndf += 2;
    }
    const int64_t slot = base + k;
    if (out.smoothed)
      for (int i = 0; i < 6; ++i) out.smoothed[6 * slot + i] = xs[i];
    if (out.smoothedCov) storeSym(out.smoothedCov + 36 * slot, Cs);
    if (out.filtered)
      for (int i = 0; i < 6; ++i) out.filtered[6 * slot + i] = xf[i];
    if (out.filteredCov) storeSym(out.filteredCov + 36 * slot, Cf);
    if (out.predicted)
      for (int i = 0; i < 6; ++i) out.predicted[6 * slot + i] = sd(sc, k, kPred + i, g);
    if (out.predictedCov) {
      double P[21];
      for (int i = 0; i < 21; ++i) P[i] = sd(sc, k, kPCov + i, g);
      storeSym(out.predictedCov + 36 * slot, P);
    }
    if (out.jacobian)
      for (int i = 0; i < 36; ++i) out.jacobian[36 * slot + i] = sd(sc, k, kJac + i, g);
    if (out.stateSurface) out.stateSurface[slot] = si(sc, k, kISurf, g);
    if (out.stateFlags) out.stateFlags[slot] = flags;
    if (out.statePath) out.statePath[slot] = sd(sc, k, kPath, g);
    if (out.stateChi2) out.stateChi2[slot] = c2;
#pragma unroll
    for (int i = 0; i < 6; ++i) xsN[i] = xs[i];
#pragma unroll
    for (int i = 0; i < 21; ++i) CsN[i] = Cs[i];
  }
  if (!valid) return;
  if (out.status) out.status[t] = status;
  if (out.nStates) out.nStates[t] = status ? 0 : tip + 1;
  if (out.chi2) out.chi2[t] = status ? 0.f : chi2sum;
  if (out.ndf) out.ndf[t] = status ? 0 : ndf;
  if (out.nMeas) out.nMeas[t] = status ? 0 : nMeasT;
  if (out.nHoles) out.nHoles[t] = status ? 0 : nHoles;
  if (out.nOutliers) out.nOutliers[t] = status ? 0 : nOut;
}

}  // namespace kf
