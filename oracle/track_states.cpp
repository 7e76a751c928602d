// track_states.cpp -- CPU restatement of the per-state columns of
// RootTrackStatesWriter (Examples/Io/Root/src/RootTrackStatesWriter.cpp:
// 364-735) and of calculateUnbiasedParametersCovariance
// (Core/src/Utilities/TrackHelpersUnbiased.cpp:21-54).
// TEST INFRASTRUCTURE ONLY (see kf_oracle.hpp): the checker of the HIP
// track-state table, never part of the product.
//
// The writer pushes per-state values into per-track vectors, some only for
// states with a measurement.  This restatement fills one dense row per state
// slot in the column layout of acts_kf_track_states_out, with NaN where the
// writer pushes NaN and where it pushes nothing; the host-side writer
// rebuilds the ragged vectors from the row and the state type.
#include <cfloat>
#include <cmath>
#include <limits>

#include "kf_oracle.hpp"

namespace acts_oracle {

namespace {

const double kNaN = std::numeric_limits<double>::quiet_NaN();

// Acts::clampValue<float, double> (Utilities/Helpers.hpp:170-188)
float clampValue(double value) {
  if (std::isinf(value))
    return value > 0 ? std::numeric_limits<float>::infinity()
                     : -std::numeric_limits<float>::infinity();
  if (std::isnan(value)) return std::numeric_limits<float>::quiet_NaN();
  const double lo = static_cast<double>(std::numeric_limits<float>::lowest());
  const double hi = static_cast<double>(std::numeric_limits<float>::max());
  return static_cast<float>(value < lo ? lo : (hi < value ? hi : value));
}

// detail::difference_periodic (Utilities/detail/periodic.hpp:38-50)
double differencePeriodic(double lhs, double rhs, double range) {
  double delta = std::fmod(lhs - rhs, range);
  if ((2 * delta) < -range) {
    delta += range;
  } else if (range <= (2 * delta)) {
    delta -= range;
  }
  return delta;
}

// VectorHelpers::eta (Utilities/VectorHelpers.hpp:119-127), with perp()
// the transverse norm (:74-86)
double eta(const V3& v) {
  if (v[0] == 0. && v[1] == 0.)
    return std::copysign(std::numeric_limits<double>::infinity(), v[2]);
  return std::asinh(v[2] / perp(v));
}

struct Row {
  float* cols;
  int32_t* ints;
  int64_t stride, slot;
  void put(int c, double v) const { cols[c * stride + slot] = clampValue(v); }
  void putRaw(int c, float v) const { cols[c * stride + slot] = v; }
  void puti(int c, int32_t v) const { ints[c * stride + slot] = v; }
};

template <int R, int C>
Mat<R, C> add(const Mat<R, C>& A, const Mat<R, C>& B) {
  Mat<R, C> o;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) o.a[i][j] = A.a[i][j] + B.a[i][j];
  return o;
}
template <int R, int C>
Mat<R, C> sub(const Mat<R, C>& A, const Mat<R, C>& B) {
  Mat<R, C> o;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) o.a[i][j] = A.a[i][j] - B.a[i][j];
  return o;
}

// H: the N x 6 projector of a measurement on the bound parameters idx
template <int N>
Mat<N, 6> projector(const int idx[2]) {
  Mat<N, 6> H = Mat<N, 6>::zero();
  for (int i = 0; i < N; ++i) H.a[i][idx[i]] = 1;
  return H;
}

// Eigen's closed-form 2x2 inverse (invdet = 1 / det, det = a d - c b)
Mat<2, 2> inverse2(const Mat<2, 2>& m) {
  const double det = m.a[0][0] * m.a[1][1] - m.a[1][0] * m.a[0][1];
  const double invdet = 1.0 / det;
  Mat<2, 2> r;
  r.a[0][0] = m.a[1][1] * invdet;
  r.a[1][0] = -m.a[1][0] * invdet;
  r.a[0][1] = -m.a[0][1] * invdet;
  r.a[1][1] = m.a[0][0] * invdet;
  return r;
}

// TrackHelpersUnbiased.cpp:37-52 for a 2D measurement
inline Mat<1, 1> inverseN(const Mat<1, 1>& m) {
  Mat<1, 1> r;
  r.a[0][0] = 1.0 / m.a[0][0];
  return r;
}
inline Mat<2, 2> inverseN(const Mat<2, 2>& m) { return inverse2(m); }

// calculateUnbiasedParametersCovariance (TrackHelpersUnbiased.cpp:37-52)
template <int N>
void unbiased(const Vec6& s, const M6& C, const int idx[2], const Mat<N, 1>& m,
              const Mat<N, N>& V, Vec6& su, M6& Cu) {
  const Mat<N, 6> H = projector<N>(idx);
  const Mat<6, N> Ht = transpose(H);
  // the products keep the projector's zero terms, as Eigen's do
  const Mat<6, N> K = mul(mul(C, Ht), inverseN(sub(mul(mul(H, C), Ht), V)));
  su = add(s, mul(K, sub(m, mul(H, s))));
  Cu = sub(C, mul(mul(K, H), C));
}

// the measurement of a state (dimension N = ms.dim): values and covariance
template <int N>
void measMats(const Meas& ms, Mat<N, 1>& m, Mat<N, N>& V) {
  for (int i = 0; i < N; ++i) m.a[i][0] = ms.v[i];
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) V.a[i][j] = ms.V[N * i + j];
}

// local hit residual against the prediction (RootTrackStatesWriter.cpp:
// 677-722): res = m - H x, resCov = V + H C H^T; the x columns take the
// first measured component, the y columns the second (NaN for N = 1)
template <int N>
void hitResiduals(const Row& row, const Meas& ms, const Vec6& x, const M6& C) {
  Mat<N, 1> m;
  Mat<N, N> V;
  measMats<N>(ms, m, V);
  const Mat<N, 6> H = projector<N>(ms.idx);
  const Mat<N, N> resCov = add(V, mul(mul(H, C), transpose(H)));
  const Mat<N, 1> res = sub(m, mul(H, x));
  row.put(ACTS_KF_TS_RES_X_HIT + 0, res.a[0][0]);
  row.put(ACTS_KF_TS_RES_X_HIT + 2, V.a[0][0] >= 0 ? std::sqrt(V.a[0][0]) : kNaN);
  row.put(ACTS_KF_TS_RES_X_HIT + 4,
          resCov.a[0][0] > 0 ? res.a[0][0] / std::sqrt(resCov.a[0][0]) : kNaN);
  if (N >= 2) {
    row.put(ACTS_KF_TS_RES_X_HIT + 1, res.a[N - 1][0]);
    row.put(ACTS_KF_TS_RES_X_HIT + 3,
            V.a[N - 1][N - 1] >= 0 ? std::sqrt(V.a[N - 1][N - 1]) : kNaN);
    row.put(ACTS_KF_TS_RES_X_HIT + 5, resCov.a[N - 1][N - 1] > 0
                                          ? res.a[N - 1][0] / std::sqrt(resCov.a[N - 1][N - 1])
                                          : kNaN);
  } else {
    row.put(ACTS_KF_TS_RES_X_HIT + 1, kNaN);
    row.put(ACTS_KF_TS_RES_X_HIT + 3, kNaN);
    row.put(ACTS_KF_TS_RES_X_HIT + 5, kNaN);
  }
}

template <int N>
void unbiasedOf(const Meas& ms, const Vec6& s, const M6& C, Vec6& su, M6& Cu) {
  Mat<N, 1> m;
  Mat<N, N> V;
  measMats<N>(ms, m, V);
  unbiased<N>(s, C, ms.idx, m, V, su, Cu);
}

// the values, errors, global position and momentum, residuals and pulls of
// one parameter set (RootTrackStatesWriter.cpp:576-676)
void fillParams(const Row& row, int base, const Surface& surface, const Vec6& x, const M6& C,
                const Vec6* truth) {
  double err[6];
  for (int i = 0; i < 6; ++i) {
    row.put(base + i, x.a[i][0]);
    const double variance = C.a[i][i];
    err[i] = variance >= 0 ? std::sqrt(variance) : kNaN;
    row.put(base + 12 + i, err[i]);
  }
  const Vec8 f = boundToFree(surface, x);
  row.put(base + 24, f.a[0][0]);
  row.put(base + 25, f.a[1][0]);
  row.put(base + 26, f.a[2][0]);
  const double p = std::abs(1 / f.a[7][0]);
  row.put(base + 27, p * f.a[4][0]);
  row.put(base + 28, p * f.a[5][0]);
  row.put(base + 29, p * f.a[6][0]);
  row.put(base + 30, eta(v3(f.a[4][0], f.a[5][0], f.a[6][0])));
  row.put(base + 31, p * std::hypot(f.a[4][0], f.a[5][0]));
  for (int i = 0; i < 6; ++i) {
    double res = kNaN, pull = kNaN;
    if (truth) {
      res = x.a[i][0] - truth->a[i][0];
      if (i == 2) res = differencePeriodic(x.a[2][0], truth->a[2][0], 2 * M_PI);
      pull = (!std::isnan(err[i]) && err[i] > 0) ? res / err[i] : kNaN;
    }
    row.put(base + 6 + i, res);
    row.put(base + 18 + i, pull);
  }
}

Vec6 loadVec(const double* p) {
  Vec6 v;
  for (int i = 0; i < 6; ++i) v.a[i][0] = p[i];
  return v;
}
M6 loadMat(const double* p) {
  M6 m;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) m.a[i][j] = p[6 * i + j];
  return m;
}

void emptyRow(const Row& row) {
  row.puti(ACTS_KF_TS_STATE_TYPE, -1);
  row.puti(ACTS_KF_TS_DIM_HIT, 0);
  row.puti(ACTS_KF_TS_HAS_PARAMS, 0);
  for (int c = 0; c < ACTS_KF_TS_NFLOAT; ++c) row.put(c, kNaN);
}

}  // namespace

void trackStatesTable(const Geometry& geo, const acts_kf_batch_in& in,
                      const acts_kf_batch_out& f, const double* truthMeas,
                      const acts_kf_track_states_out& out) {
  const int64_t n = in.n_tracks;
  const int64_t nMeas = in.meas_offsets[n];
  for (int64_t slot = 0; slot < out.n_slots; ++slot) {
    Row row{out.columns, out.ints, out.n_slots, slot};
    // the slot's track
    int64_t t = 0;
    {
      int64_t lo = 0, hi = n;
      while (hi - lo > 1) {
        const int64_t mid = (lo + hi) / 2;
        if (in.seq_offsets[mid] <= slot) lo = mid; else hi = mid;
      }
      t = lo;
    }
    const int64_t k = slot - in.seq_offsets[t];
    const bool inSeq = slot < in.seq_offsets[n] && k >= 0 && slot < in.seq_offsets[t + 1];
    row.puti(ACTS_KF_TS_TRACK, inSeq ? static_cast<int32_t>(t) : -1);
    if (!inSeq || f.status[t] != 0 || k >= f.n_states[t]) {
      emptyRow(row);
      continue;
    }
    const int32_t sid = f.state_surface[slot];
    const int32_t mi = f.state_measurement[slot];
    if (sid < 0 || sid >= static_cast<int32_t>(geo.surfaces.size()) || mi >= nMeas) {
      emptyRow(row);
      continue;
    }
    const Surface& surface = geo.surfaces[sid];
    const uint32_t flags = f.state_flags[slot];
    // getStateType (RootTrackStatesWriter.cpp:293-308)
    int type = 4;
    if (flags & ACTS_KF_STATE_IS_OUTLIER) type = 1;
    else if (flags & ACTS_KF_STATE_HAS_MEASUREMENT) type = 0;
    else if (flags & ACTS_KF_STATE_IS_HOLE) type = 2;
    else if (flags & ACTS_KF_STATE_HAS_MATERIAL) type = 3;
    row.puti(ACTS_KF_TS_STATE_TYPE, type);
    const bool hasSourceLink = mi >= 0;
    row.putRaw(ACTS_KF_TS_CHI2, f.state_chi2[slot]);
    row.putRaw(ACTS_KF_TS_PATH_LENGTH, static_cast<float>(f.state_path[slot]));

    // truth at the hit (RootTrackStatesWriter.cpp:390-478): the truth
    // particle's bound state on the surface stands for the averaged sim hits
    Vec6 truth;
    const bool haveTruth = hasSourceLink && truthMeas != nullptr;
    Meas ms;
    if (haveTruth) {
      truth = loadVec(truthMeas + 6 * static_cast<int64_t>(mi));
      const Vec8 tf = boundToFree(surface, truth);
      const V3 tpos = v3(tf.a[0][0], tf.a[1][0], tf.a[2][0]);
      row.put(ACTS_KF_TS_T_X + 0, tpos[0]);
      row.put(ACTS_KF_TS_T_X + 1, tpos[1]);
      row.put(ACTS_KF_TS_T_X + 2, tpos[2]);
      row.put(ACTS_KF_TS_T_X + 3, perp(tpos));
      row.put(ACTS_KF_TS_T_X + 4, tf.a[4][0]);
      row.put(ACTS_KF_TS_T_X + 5, tf.a[5][0]);
      row.put(ACTS_KF_TS_T_X + 6, tf.a[6][0]);
      for (int i = 0; i < 6; ++i) row.put(ACTS_KF_TS_T_PARAMS + i, truth.a[i][0]);
    } else {
      for (int c = ACTS_KF_TS_T_X; c < ACTS_KF_TS_L_X_HIT; ++c) row.put(c, kNaN);
    }
    if (hasSourceLink) {
      const int64_t g = mi;
      ms.dim = in.meas_dims ? in.meas_dims[g] : 2;
      if (in.meas_indices) {
        ms.idx[0] = in.meas_indices[2 * g];
        ms.idx[1] = in.meas_indices[2 * g + 1];
      }
      for (int i = 0; i < 2; ++i) ms.v[i] = in.meas_values[2 * g + i];
      for (int i = 0; i < 4; ++i) ms.V[i] = in.meas_cov[4 * g + i];
      // the measurement expanded into the bound space: its loc0 and loc1
      // (0 where not measured)
      double loc[2] = {0.0, 0.0};
      for (int i = 0; i < ms.dim; ++i)
        if (ms.idx[i] < 2) loc[ms.idx[i]] = ms.v[i];
      // the plane, disc and cylinder localToGlobal ignore the direction
      const V3 global = surface.localToGlobal(loc[0], loc[1]);
      row.put(ACTS_KF_TS_L_X_HIT + 0, loc[0]);
      row.put(ACTS_KF_TS_L_X_HIT + 1, loc[1]);
      row.put(ACTS_KF_TS_L_X_HIT + 2, global[0]);
      row.put(ACTS_KF_TS_L_X_HIT + 3, global[1]);
      row.put(ACTS_KF_TS_L_X_HIT + 4, global[2]);
    } else {
      for (int c = ACTS_KF_TS_L_X_HIT; c < ACTS_KF_TS_RES_X_HIT; ++c) row.put(c, kNaN);
    }

    const double* par[3] = {f.predicted, f.filtered, f.smoothed};
    const double* cov[3] = {f.predicted_cov, f.filtered_cov, f.smoothed_cov};
    Vec6 smoothed;
    M6 smoothedCov;
    for (int p = 0; p < 3; ++p) {
      const Vec6 x = loadVec(par[p] + 6 * slot);
      const M6 C = loadMat(cov[p] + 36 * slot);
      fillParams(row, ACTS_KF_TS_PARAMS(p), surface, x, C, haveTruth ? &truth : nullptr);
      if (p == ACTS_KF_TS_PREDICTED) {
        // local hit residual (RootTrackStatesWriter.cpp:677-722)
        if (hasSourceLink) {
          if (ms.dim == 1)
            hitResiduals<1>(row, ms, x, C);
          else
            hitResiduals<2>(row, ms, x, C);
        } else {
          for (int c = ACTS_KF_TS_RES_X_HIT; c < ACTS_KF_TS_PARAMS(0); ++c) row.put(c, kNaN);
        }
      }
      if (p == ACTS_KF_TS_SMOOTHED) {
        smoothed = x;
        smoothedCov = C;
      }
    }
    // unbiased parameters: smoothed, with projector and calibrated data
    if (hasSourceLink) {
      Vec6 su;
      M6 Cu;
      if (ms.dim == 1)
        unbiasedOf<1>(ms, smoothed, smoothedCov, su, Cu);
      else
        unbiasedOf<2>(ms, smoothed, smoothedCov, su, Cu);
      fillParams(row, ACTS_KF_TS_PARAMS(ACTS_KF_TS_UNBIASED), surface, su, Cu,
                 haveTruth ? &truth : nullptr);
    } else {
      for (int c = ACTS_KF_TS_PARAMS(ACTS_KF_TS_UNBIASED); c < ACTS_KF_TS_NFLOAT; ++c)
        row.put(c, kNaN);
    }
    row.puti(ACTS_KF_TS_DIM_HIT, hasSourceLink ? ms.dim : 0);
    row.puti(ACTS_KF_TS_HAS_PARAMS, hasSourceLink ? 0xF : 0x7);
  }
}

}  // namespace acts_oracle
