// Track-state table: the per-state columns of RootTrackStatesWriter
// (Examples/Io/Root/src/RootTrackStatesWriter.cpp:364-735) computed on the
// device from a fitted batch, one thread per state slot, column-major
// float output so that every column store of a wavefront is one coalesced
// 256-byte line.  The work per state is ~1.1 KB of reads (three parameter
// vectors and covariances, the measurement, the truth) and 0.6 KB of
// writes against ~600 FLOP: the kernel is HBM-bound.
#pragma once

#include <cfloat>

#include "kf_device.hpp"

namespace kf {

struct TsArgs {
  int64_t nTracks;
  int64_t nSlots;  // column stride and number of rows written
  int32_t nSurfaces;
  int32_t pad;
  const int64_t* seqOff;
  const int64_t* measOff;
  const double* measVal;
  const double* measCov;
  const uint8_t* measDims;  // null: all 2D on (loc0, loc1)
  const uint8_t* measIdx;
  const int32_t* status;
  const int32_t* nStates;
  const int32_t* stateSurf;
  const uint32_t* stateFlags;
  const double* statePath;
  const float* stateChi2;
  const int32_t* stateMeas;
  const double* par[3];  // predicted, filtered, smoothed [slots][6]
  const double* cov[3];  // [slots][36]
  const double* truth;   // [m][6] or null
  const DevSurface* geo;
  float* cols;
  int32_t* ints;
};

// clampValue<float>(double) (Helpers.hpp:170-188): infinities and NaN kept,
// finite values clamped to the float range, then rounded to nearest
// Branch-free: the conversion keeps infinities and NaN and rounds finite
// values to nearest; a finite value that rounds past the float range is
// the clamped one, +-FLT_MAX (values that round down to FLT_MAX are
// FLT_MAX either way).
KF_HD float clampF(double v) {
  const float f = static_cast<float>(v);
  return (isinf(f) && !isinf(v)) ? copysignf(FLT_MAX, f) : f;
}

// detail::difference_periodic (periodic.hpp:38-50)
KF_HD double differencePeriodic(double lhs, double rhs, double range) {
  double delta = fmod(lhs - rhs, range);
  if ((2 * delta) < -range) {
    delta += range;
  } else if (range <= (2 * delta)) {
    delta -= range;
  }
  return delta;
}

// VectorHelpers::eta (VectorHelpers.hpp:119-127)
KF_HD double etaOf(const V3& v) {
  if (v.x == 0. && v.y == 0.) return copysign(static_cast<double>(INFINITY), v.z);
  return asinh(v.z / perp(v));
}

// v[k] for a run-time k in 0..5, by selects (a register array indexed at
// run time would be moved to scratch memory)
KF_HD double pick6(const double* v, int k) {
  double r = v[0];
#pragma unroll
  for (int i = 1; i < 6; ++i) r = (k == i) ? v[i] : r;
  return r;
}

// calculateUnbiasedParametersCovariance (TrackHelpersUnbiased.cpp:37-52)
// for a measurement of dimension n on the bound parameters a (, b):
// K = C H^T (H C H^T - V)^-1, s + K (m - H s), C - K H C, with Eigen's
// closed-form 1x1 / 2x2 inverse.  The table needs the parameters and the
// diagonal of the covariance; C (row-major 6x6) is read in the measured
// rows and columns and on the diagonal.
KF_HD void unbiasedState(const double* s, const double* C, int n, int a, int b,
                         const double* m, const double* V, double* su, double* du) {
  const bool two = n == 2;
  const double S00 = C[7 * a] - V[0];
  const double S01 = two ? C[6 * a + b] - V[1] : 0.0;
  const double S10 = two ? C[6 * b + a] - V[2] : 0.0;
  const double S11 = two ? C[7 * b] - V[3] : 1.0;
  double Si[2][2];
  if (two) {
    const double det = S00 * S11 - S10 * S01;
    const double invdet = 1.0 / det;
    Si[0][0] = S11 * invdet;
    Si[0][1] = -S01 * invdet;
    Si[1][0] = -S10 * invdet;
    Si[1][1] = S00 * invdet;
  } else {
    Si[0][0] = 1.0 / S00;
    Si[0][1] = Si[1][0] = Si[1][1] = 0.0;
  }
  const double r0 = m[0] - pick6(s, a), r1 = two ? m[1] - pick6(s, b) : 0.0;
  for (int i = 0; i < 6; ++i) {
    double K0 = C[6 * i + a] * Si[0][0], K1 = 0.0;
    if (two) {
      K0 = K0 + C[6 * i + b] * Si[1][0];
      K1 = C[6 * i + a] * Si[0][1] + C[6 * i + b] * Si[1][1];
    }
    su[i] = s[i] + (two ? K0 * r0 + K1 * r1 : K0 * r0);
    du[i] = C[7 * i] - (two ? K0 * C[6 * a + i] + K1 * C[6 * b + i] : K0 * C[6 * a + i]);
  }
}

// The 32 columns of one parameter set (values, residuals, errors, pulls,
// global position, momentum, eta, pT) for a state on surface s.  `truth`
// is null when the state has no measurement or no truth was given: the
// residuals and pulls are then NaN.
template <typename Store>
KF_HD void paramColumns(const DevSurface& s, const double* x, const double* diag,
                        const double* truth, Store&& put, int base) {
  double err[6];
  for (int i = 0; i < 6; ++i) {
    const double var = diag[i];
    err[i] = var >= 0 ? sqrt(var) : static_cast<double>(NAN);
    put(base + i, clampF(x[i]));
    put(base + 12 + i, clampF(err[i]));
  }
  // transformBoundToFreeParameters: direction from (phi, theta), position
  // from the surface's localToGlobal
  double sphi, cphi, sth, cth;
  sincos(x[2], &sphi, &cphi);
  sincos(x[3], &sth, &cth);
  const V3 dir = mk(cphi * sth, sphi * sth, cth);
  const V3 pos = localToGlobal(s, x[0], x[1]);
  const double p = fabs(1 / x[4]);
  put(base + 24, clampF(pos.x));
  put(base + 25, clampF(pos.y));
  put(base + 26, clampF(pos.z));
  put(base + 27, clampF(p * dir.x));
  put(base + 28, clampF(p * dir.y));
  put(base + 29, clampF(p * dir.z));
  put(base + 30, clampF(etaOf(dir)));
  put(base + 31, clampF(p * hypot(dir.x, dir.y)));
  for (int i = 0; i < 6; ++i) {
    double res = static_cast<double>(NAN), pull = static_cast<double>(NAN);
    if (truth) {
      res = i == 2 ? differencePeriodic(x[2], truth[2], 2 * kPi) : x[i] - truth[i];
      pull = (!isnan(err[i]) && err[i] > 0) ? res / err[i] : static_cast<double>(NAN);
    }
    put(base + 6 + i, clampF(res));
    put(base + 18 + i, clampF(pull));
  }
}

constexpr int kTsBlock = 256;

// One row of the table, for slot `slot`: everything RootTrackStatesWriter
// derives from one track state (the body of its state loop,
// RootTrackStatesWriter.cpp:364-735).  The row is computed in FP64 and
// stored as float: the cost is the FP64 transcendental and division work
// of four parameter sets (~5 k lane instructions per state), not the
// 1.8 KB it moves.
// Sub: the batch has measurement subspaces (dimension and indices read per
// measurement); otherwise every measurement is 2D on (loc0, loc1), known at
// compile time.
template <bool Sub>
__device__ inline void trackStateRow(const TsArgs& a, int64_t slot) {
  const int64_t nSeq = a.seqOff[a.nTracks];
  const bool inTable = slot < a.nSlots;
  auto put = [&](int c, float v) { a.cols[c * a.nSlots + slot] = v; };
  auto puti = [&](int c, int32_t v) { a.ints[c * a.nSlots + slot] = v; };
  // the slot's track: the last t with seqOff[t] <= slot
  int64_t t = -1;
  if (inTable && slot < nSeq) {
    int64_t lo = 0, hi = a.nTracks;
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (a.seqOff[mid] <= slot) lo = mid; else hi = mid;
    }
    t = lo;
  }
  const int64_t k = t < 0 ? -1 : slot - a.seqOff[t];
  const bool inSeq = t >= 0 && k >= 0 && slot < a.seqOff[t + 1];
  bool ok = inSeq && a.status[t] == 0 && k < a.nStates[t];
  // indices taken from the fit's outputs are checked before any lookup
  const int32_t sid = ok ? a.stateSurf[slot] : -1;
  const int32_t mi = ok ? a.stateMeas[slot] : -1;
  ok = ok && sid >= 0 && sid < a.nSurfaces && mi < a.measOff[a.nTracks];
  const bool link = ok && mi >= 0;
  if (inTable) puti(ACTS_KF_TS_TRACK, inSeq ? static_cast<int32_t>(t) : -1);
  if (inTable && !ok) {
    puti(ACTS_KF_TS_STATE_TYPE, -1);
    puti(ACTS_KF_TS_DIM_HIT, 0);
    puti(ACTS_KF_TS_HAS_PARAMS, 0);
    for (int c = 0; c < ACTS_KF_TS_NFLOAT; ++c) put(c, NAN);
  }
  const DevSurface& s = a.geo[ok ? sid : 0];
  const double* truth = (link && a.truth) ? a.truth + 6 * static_cast<int64_t>(mi) : nullptr;
  double m[2] = {0, 0}, V[4] = {0, 0, 0, 0};
  int mdim = 2, ma = 0, mb = 1;
  if constexpr (Sub) {
    if (link && a.measDims) mdim = a.measDims[mi];
    if (link && a.measIdx) {
      ma = a.measIdx[2 * static_cast<int64_t>(mi)];
      mb = a.measIdx[2 * static_cast<int64_t>(mi) + 1];
    }
    if (mdim == 1) mb = ma;
  }
  if (ok) {
    const uint32_t flags = a.stateFlags[slot];
    // RootTrackStatesWriter::getStateType (RootTrackStatesWriter.cpp:293-308)
    const int32_t type = (flags & ACTS_KF_STATE_IS_OUTLIER)       ? 1
                         : (flags & ACTS_KF_STATE_HAS_MEASUREMENT) ? 0
                         : (flags & ACTS_KF_STATE_IS_HOLE)         ? 2
                         : (flags & ACTS_KF_STATE_HAS_MATERIAL)    ? 3
                                                                   : 4;
    puti(ACTS_KF_TS_STATE_TYPE, type);
    puti(ACTS_KF_TS_DIM_HIT, link ? mdim : 0);
    puti(ACTS_KF_TS_HAS_PARAMS, link ? 0xF : 0x7);
    put(ACTS_KF_TS_CHI2, a.stateChi2[slot]);
    put(ACTS_KF_TS_PATH_LENGTH, static_cast<float>(a.statePath[slot]));
    // truth at the hit: the truth particle's bound state on the surface
    if (truth) {
      double sphi, cphi, sth, cth;
      sincos(truth[2], &sphi, &cphi);
      sincos(truth[3], &sth, &cth);
      const V3 tdir = mk(cphi * sth, sphi * sth, cth);
      const V3 tpos = localToGlobal(s, truth[0], truth[1]);
      put(ACTS_KF_TS_T_X + 0, clampF(tpos.x));
      put(ACTS_KF_TS_T_X + 1, clampF(tpos.y));
      put(ACTS_KF_TS_T_X + 2, clampF(tpos.z));
      put(ACTS_KF_TS_T_X + 3, clampF(perp(tpos)));
      put(ACTS_KF_TS_T_X + 4, clampF(tdir.x));
      put(ACTS_KF_TS_T_X + 5, clampF(tdir.y));
      put(ACTS_KF_TS_T_X + 6, clampF(tdir.z));
      for (int i = 0; i < 6; ++i) put(ACTS_KF_TS_T_PARAMS + i, clampF(truth[i]));
    } else {
      for (int c = ACTS_KF_TS_T_X; c < ACTS_KF_TS_L_X_HIT; ++c) put(c, NAN);
    }
    // the hit: local and global position (plane, disc and cylinder
    // localToGlobal do not depend on the direction)
    if (link) {
      for (int i = 0; i < 2; ++i) m[i] = a.measVal[2 * static_cast<int64_t>(mi) + i];
      for (int i = 0; i < 4; ++i) V[i] = a.measCov[4 * static_cast<int64_t>(mi) + i];
      // the measurement expanded into the bound space: its loc0 and loc1
      // (0 where not measured)
      const double l0 = (ma == 0) ? m[0] : ((mdim == 2 && mb == 0) ? m[1] : 0.0);
      const double l1 = (ma == 1) ? m[0] : ((mdim == 2 && mb == 1) ? m[1] : 0.0);
      const V3 g = localToGlobal(s, l0, l1);
      put(ACTS_KF_TS_L_X_HIT + 0, clampF(l0));
      put(ACTS_KF_TS_L_X_HIT + 1, clampF(l1));
      put(ACTS_KF_TS_L_X_HIT + 2, clampF(g.x));
      put(ACTS_KF_TS_L_X_HIT + 3, clampF(g.y));
      put(ACTS_KF_TS_L_X_HIT + 4, clampF(g.z));
    } else {
      for (int c = ACTS_KF_TS_L_X_HIT; c < ACTS_KF_TS_RES_X_HIT; ++c) put(c, NAN);
    }
  }

  if (!ok) return;
  // the predicted, filtered and smoothed sets; the unbiased set is formed
  // from the smoothed one (not unrolled: one copy of the transcendental
  // calls)
  double x[6], d[6], xu[6], du[6];
#pragma unroll 1
  for (int p = 0; p < 3; ++p) {
    const double* C = a.cov[p] + 36 * slot;
    const double* xp = a.par[p] + 6 * slot;
    for (int i = 0; i < 6; ++i) x[i] = xp[i];
    for (int i = 0; i < 6; ++i) d[i] = C[7 * i];
    if (p == ACTS_KF_TS_SMOOTHED && link) unbiasedState(x, C, mdim, ma, mb, m, V, xu, du);
    paramColumns(s, x, d, truth, put, ACTS_KF_TS_PARAMS(p));
    if (p == ACTS_KF_TS_PREDICTED) {
      // hit residual against the prediction: res = m - H x, resCov = V + H C H^T
      if (link) {
        // the x columns take the first measured component, the y columns
        // the second (NaN for a 1D measurement)
        const double nan = static_cast<double>(NAN);
        const double resX = m[0] - pick6(x, ma);
        const double errX = V[0] >= 0 ? sqrt(V[0]) : nan;
        const double rcX = V[0] + pick6(d, ma);
        const double pullX = rcX > 0 ? resX / sqrt(rcX) : nan;
        double resY = nan, errY = nan, pullY = nan;
        if (mdim == 2) {
          resY = m[1] - pick6(x, mb);
          errY = V[3] >= 0 ? sqrt(V[3]) : nan;
          const double rcY = V[3] + pick6(d, mb);
          pullY = rcY > 0 ? resY / sqrt(rcY) : nan;
        }
        put(ACTS_KF_TS_RES_X_HIT + 0, clampF(resX));
        put(ACTS_KF_TS_RES_X_HIT + 1, clampF(resY));
        put(ACTS_KF_TS_RES_X_HIT + 2, clampF(errX));
        put(ACTS_KF_TS_RES_X_HIT + 3, clampF(errY));
        put(ACTS_KF_TS_RES_X_HIT + 4, clampF(pullX));
        put(ACTS_KF_TS_RES_X_HIT + 5, clampF(pullY));
      } else {
        for (int c = ACTS_KF_TS_RES_X_HIT; c < ACTS_KF_TS_PARAMS(0); ++c) put(c, NAN);
      }
    }
  }
  if (link) {
    paramColumns(s, xu, du, truth, put, ACTS_KF_TS_PARAMS(ACTS_KF_TS_UNBIASED));
  } else {
    for (int c = ACTS_KF_TS_PARAMS(ACTS_KF_TS_UNBIASED); c < ACTS_KF_TS_NFLOAT; ++c) put(c, NAN);
  }
}

}  // namespace kf
