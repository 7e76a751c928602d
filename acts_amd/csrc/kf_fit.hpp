// kf_fit.hpp -- per-track forward filter + RTS smoother on the device.
#pragma once

#include "kf_mixed.hpp"

namespace kf {

struct DevOut {
  int32_t* status;
  int32_t* nStates;
  float* chi2;
  int32_t* ndf;
  int32_t* nMeas;
  int32_t* nHoles;
  int32_t* nOutliers;
  int32_t* nSteps;
  int32_t* nTrials;
  int32_t* stateSurface;
  uint32_t* stateFlags;
  double* statePath;
  float* stateChi2;
  double* predicted;
  double* predictedCov;
  double* filtered;
  double* filteredCov;
  double* smoothed;
  double* smoothedCov;
  double* jacobian;
  double* referenceParams;
  double* referenceCov;
  int32_t* stateMeas;
  int32_t* reversed;
};

// ---------------------------------------------------------------------------
// detail::transportCovarianceToBound (CovarianceEngine.cpp:102-165) with
// boundToBoundTransportJacobian (JacobianEngine.cpp:19-45) and
// reinitializeJacobians (JacobianEngine.cpp:122-141).  The dense chain
// F2B * (I + d s) * Jt * B2F is evaluated in the reference's association
// order with its structural zeros skipped.  `onSurface` is the result of
// globalToLocal at the current position (reinitialization keeps the old
// jacToGlobal when it fails, as the reference ignores that Result).
// ---------------------------------------------------------------------------
// With Sympy the jacobian is F2B * (I + d s) * M, M the bound-to-free
// jacobian the SympyStepper carries (sympy::boundToBoundTransportJacobian,
// SympyJacobianEngine.cpp:21-45), in the same association order with its
// structural zeros skipped.
template <bool Sympy = false>
__device__ inline void transportToBound(Track& T, const DevSurface& s,
                                        bool onSurface, double J[36]) {
  const V3 dir = T.dir;
  double b2f[6], f2b[6];
  positionJacobians(s, T.pos, b2f, f2b);
  V3 fz;
  if (s.type == ACTS_KF_SURFACE_CYLINDER)
    fz = cylinderNormal(s, T.pos);
  else
    fz = col(s.R, 2);
  // freeToPathDerivative (Surface.cpp:292-307)
  const double dz = dot(fz, dir);
  const double sv[3] = {(-1.0 * fz.x) / dz, (-1.0 * fz.y) / dz, (-1.0 * fz.z) / dz};
  // direction rows of freeToBoundJacobian (JacobianHelpers.hpp:52-64)
  const double cosTheta = dir.z;
  const double sinTheta = cathetus(1, dir.z);
  const double inv = 1. / sinTheta;
  const double cosPhi = dir.x * inv, sinPhi = dir.y * inv;
  const double invSinTheta = 1. / sinTheta;
  const double F24 = -sinPhi * invSinTheta, F25 = cosPhi * invSinTheta;
  const double F34 = cosPhi * cosTheta, F35 = sinPhi * cosTheta, F36 = -sinTheta;

  // A = I + derivative * freeToPath, columns 0..2 (A(7, j) is zero)
  double A[7][3];
#pragma unroll
  for (int k = 0; k < 7; ++k)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double ds = T.der[k] * sv[j];
      A[k][j] = (k == j) ? 1.0 + ds : ds;
    }
  // A1 = F2B * A, columns 0..2 (columns 3..7 equal F2B)
  double A1[6][3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    A1[0][j] = (f2b[0] * A[0][j] + f2b[1] * A[1][j]) + f2b[2] * A[2][j];
    A1[1][j] = (f2b[3] * A[0][j] + f2b[4] * A[1][j]) + f2b[5] * A[2][j];
    A1[2][j] = F24 * A[4][j] + F25 * A[5][j];
    A1[3][j] = (F34 * A[4][j] + F35 * A[5][j]) + F36 * A[6][j];
    A1[4][j] = 0.0;
    A1[5][j] = A[3][j];
  }
  if constexpr (Sympy) {
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      const double a0 = (r == 4) ? 0.0 : A1[r][0];
      const double a1 = (r == 4) ? 0.0 : A1[r][1];
      const double a2 = (r == 4) ? 0.0 : A1[r][2];
      J[6 * r + 0] = (a0 * T.jp[0] + a1 * T.jp[2]) + a2 * T.jp[4];
      J[6 * r + 1] = (a0 * T.jp[1] + a1 * T.jp[3]) + a2 * T.jp[5];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double m0 = T.tr[c], m1 = T.tr[3 + c], m2 = T.tr[6 + c];
        const double d0 = T.br[c], d1 = T.br[3 + c], d2 = T.br[6 + c];
        double acc;
        if (r == 4) {
          acc = (c == 2) ? 1.0 : 0.0;  // M(7, 2 + c)
        } else {
          acc = (a0 * m0 + a1 * m1) + a2 * m2;
          if (r == 2) acc = (acc + F24 * d0) + F25 * d1;
          if (r == 3) acc = ((acc + F34 * d0) + F35 * d1) + F36 * d2;
          if (r == 5 && c == 2) acc = acc + T.tr37;  // M(3, 4)
        }
        J[6 * r + 2 + c] = acc;
      }
      J[6 * r + 5] = (r == 5) ? 1.0 : 0.0;
    }
  } else {
  // A2 = A1 * Jt, columns 4..7 (columns 0..3 equal A1's)
  double A2[6][4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const double t0 = T.tr[c], t1 = T.tr[4 + c], t2 = T.tr[8 + c];
    const double r0 = T.br[c], r1 = T.br[4 + c], r2 = T.br[8 + c];
    A2[0][c] = (A1[0][0] * t0 + A1[0][1] * t1) + A1[0][2] * t2;
    A2[1][c] = (A1[1][0] * t0 + A1[1][1] * t1) + A1[1][2] * t2;
    A2[2][c] = ((((A1[2][0] * t0 + A1[2][1] * t1) + A1[2][2] * t2) + F24 * r0) + F25 * r1);
    A2[3][c] = (((((A1[3][0] * t0 + A1[3][1] * t1) + A1[3][2] * t2) + F34 * r0) + F35 * r1) +
                F36 * r2);
    A2[4][c] = (c == 3) ? 1.0 : 0.0;
    double a5 = (A1[5][0] * t0 + A1[5][1] * t1) + A1[5][2] * t2;
    if (c == 3) a5 = a5 + T.tr37;
    A2[5][c] = a5;
  }
  // J = A2 * B2F
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    const double a0 = (r == 4) ? 0.0 : A1[r][0];
    const double a1 = (r == 4) ? 0.0 : A1[r][1];
    const double a2 = (r == 4) ? 0.0 : A1[r][2];
    J[6 * r + 0] = (a0 * T.jp[0] + a1 * T.jp[2]) + a2 * T.jp[4];
    J[6 * r + 1] = (a0 * T.jp[1] + a1 * T.jp[3]) + a2 * T.jp[5];
    J[6 * r + 2] = A2[r][0] * T.jd[0] + A2[r][1] * T.jd[2];
    J[6 * r + 3] = (A2[r][0] * T.jd[1] + A2[r][1] * T.jd[3]) + A2[r][2] * T.jd[4];
    J[6 * r + 4] = A2[r][3];
    J[6 * r + 5] = (r == 5) ? 1.0 : 0.0;
  }
  }
  // cov = (J * cov) * J^T on the packed symmetric covariance; row 4 of J is
  // e4 and column 5 is e5, so those products are skipped exactly.
  double JC[36];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double acc;
      if (i == 4) {
        acc = T.cov[sym(4, j)];
      } else {
        acc = J[6 * i] * T.cov[sym(0, j)];
#pragma unroll
        for (int k = 1; k < 5; ++k) acc = acc + J[6 * i + k] * T.cov[sym(k, j)];
        if (i == 5) acc = acc + T.cov[sym(5, j)];
      }
      JC[6 * i + j] = acc;
    }
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = i; j < 6; ++j) {
      double acc;
      if (j == 4) {
        acc = JC[6 * i + 4];
      } else {
        acc = JC[6 * i] * J[6 * j];
#pragma unroll
        for (int k = 1; k < 5; ++k) acc = acc + JC[6 * i + k] * J[6 * j + k];
        if (j == 5) acc = acc + JC[6 * i + 5];
      }
      T.cov[sym(i, j)] = acc;
    }
  // reinitializeJacobians
  resetTransport(T);
  if (onSurface) {
#pragma unroll
    for (int i = 0; i < 6; ++i) T.jp[i] = b2f[i];
    directionB2F(dir, T.jd);
  }
  if constexpr (Sympy) sympyResetM(T);
}

// ---------------------------------------------------------------------------
// detail::performMaterialInteraction for a slab scaled by 1
// (PointwiseMaterialInteraction.hpp:165-226, .cpp:31-87), forward direction
// ---------------------------------------------------------------------------
// propDir: the propagation direction; going backwards the momentum grows
// by the energy loss (PointwiseMaterialInteraction.hpp:188-198).
__device__ inline int materialInteraction(Track& T, const DevSurface& s,
                                          const DevSlab& slab0,
                                          const PhysConst& pc, const DevOpts& o,
                                          double propDir = 1.0) {
  const double pathCorr = pathCorrection(s, T.pos, T.dir);
  double l0, l1;
  if (!globalToLocal(s, T.pos, l0, l1)) return ACTS_KF_STATUS_NOT_ON_SURFACE;
  DevSlab slab = slab0;
  const float sc = static_cast<float>(pathCorr);
  slab.thickness *= sc;
  slab.tInX0 *= sc;
  slab.tInL0 *= sc;
  if (slab.vacuum || slab.thickness <= 0.f) return 0;
  const float qopF = static_cast<float>(T.qop);
  const double mass = o.mass;
  const double absQ = o.absQ;
  double eLoss = 0, varPhi = 0, varTheta = 0, varQop = 0;
  if (o.energyLoss) eLoss = energyLossBethe(pc, slab, o.mass, qopF, o.absQ);
  if (o.multipleScattering) {
    const double t0 = theta0(pc, slab, o.absPdg, o.mass, qopF, o.absQ);
    const double n = sqrt(dot(T.dir, T.dir));
    const double sigmaPhi = t0 * (n / perp(T.dir));
    varPhi = sigmaPhi * sigmaPhi;
    varTheta = t0 * t0;
  }
  if (o.energyLoss) {
    const double sq = energyLossLandauSigmaQOverP(pc, slab, o.mass, qopF, o.absQ);
    varQop = sq * sq;
  }
  const double momentum = copysign(absQ, T.qop) / T.qop;
  const double nextE = sqrt(mass * mass + momentum * momentum) - eLoss * propDir;
  double nextP = (mass < nextE) ? cathetus(nextE, mass) : 0;
  nextP = fmax(pc.minP, nextP);
  const float signedQ = static_cast<float>(copysign(absQ, static_cast<double>(qopF)));
  T.qop = signedQ / nextP;
  T.cov[sym(2, 2)] = fmax(0., T.cov[sym(2, 2)] + varPhi);
  T.cov[sym(3, 3)] = fmax(0., T.cov[sym(3, 3)] + varTheta);
  T.cov[sym(4, 4)] = fmax(0., T.cov[sym(4, 4)] + varQop);
  return 0;
}

// Eigen compute_inverse_size2
__device__ inline void inverse2(double a, double b, double c, double d,
                                double out[4]) {
  // matrix [[a, b], [c, d]]
  const double det = a * d - c * b;
  const double invdet = 1.0 / det;
  out[0] = d * invdet;
  out[2] = -c * invdet;
  out[1] = -b * invdet;
  out[3] = a * invdet;
}

// GainMatrixUpdaterImpl.hpp:21-91 for a 2D (loc0, loc1) measurement; P and
// C are packed symmetric.
__device__ inline int gainUpdate(const double x[6], const double P[21],
                                 const double m[2], const double V[4],
                                 bool joseph, double f[6], double C[21],
                                 double& chi2) {
  double Si[4];
  inverse2(P[sym(0, 0)] + V[0], P[sym(0, 1)] + V[1], P[sym(1, 0)] + V[2],
           P[sym(1, 1)] + V[3], Si);
  double K[12];
  bool nan = false;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    K[2 * i] = P[sym(i, 0)] * Si[0] + P[sym(i, 1)] * Si[2];
    K[2 * i + 1] = P[sym(i, 0)] * Si[1] + P[sym(i, 1)] * Si[3];
    nan = nan || isnan(K[2 * i]) || isnan(K[2 * i + 1]);
  }
  if (nan) return ACTS_KF_STATUS_UPDATE_FAILED;
  const double r0 = m[0] - x[0], r1 = m[1] - x[1];
#pragma unroll
  for (int i = 0; i < 6; ++i) f[i] = x[i] + (K[2 * i] * r0 + K[2 * i + 1] * r1);
  normalizePhiTheta(f[2], f[3]);
  if (!joseph) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const double t0 = (i == 0 ? 1.0 : 0.0) - K[2 * i];
      const double t1 = (i == 1 ? 1.0 : 0.0) - K[2 * i + 1];
#pragma unroll
      for (int j = i; j < 6; ++j) {
        double acc = t0 * P[sym(0, j)] + t1 * P[sym(1, j)];
        if (i >= 2) acc = acc + P[sym(i, j)];
        C[sym(i, j)] = acc;
      }
    }
  } else {
    double tmp[36], TP[36];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = 0; j < 6; ++j)
        tmp[6 * i + j] = (i == j ? 1.0 : 0.0) - (j < 2 ? K[2 * i + j] : 0.0);
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        double acc = tmp[6 * i] * P[sym(0, j)];
#pragma unroll
        for (int k = 1; k < 6; ++k) acc = acc + tmp[6 * i + k] * P[sym(k, j)];
        TP[6 * i + j] = acc;
      }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = i; j < 6; ++j) {
        double acc = TP[6 * i] * tmp[6 * j];
#pragma unroll
        for (int k = 1; k < 6; ++k) acc = acc + TP[6 * i + k] * tmp[6 * j + k];
        const double kv0 = K[2 * i] * V[0] + K[2 * i + 1] * V[2];
        const double kv1 = K[2 * i] * V[1] + K[2 * i + 1] * V[3];
        const double b = kv0 * K[2 * j] + kv1 * K[2 * j + 1];
        C[sym(i, j)] = acc + b;
      }
  }
  // chi2 on the filtered residual with ((I - H K) V)^-1
  const double e0 = m[0] - f[0], e1 = m[1] - f[1];
  const double i00 = 1.0 - K[0], i01 = 0.0 - K[1], i10 = 0.0 - K[2], i11 = 1.0 - K[3];
  double mi[4];
  inverse2(i00 * V[0] + i01 * V[2], i00 * V[1] + i01 * V[3],
           i10 * V[0] + i11 * V[2], i10 * V[1] + i11 * V[3], mi);
  const double row0 = e0 * mi[0] + e1 * mi[2];
  const double row1 = e0 * mi[1] + e1 * mi[3];
  chi2 = row0 * e0 + row1 * e1;
  return 0;
}

// calculatePredictedChi2 (TrackHelpers.hpp:676-694), 2D measurement
__device__ inline double predictedChi2(const double x[6], const double P[21],
                                       const double m[2], const double V[4]) {
  const double r0 = m[0] - x[0], r1 = m[1] - x[1];
  double Si[4];
  inverse2(P[sym(0, 0)] + V[0], P[sym(0, 1)] + V[1], P[sym(1, 0)] + V[2],
           P[sym(1, 1)] + V[3], Si);
  const double row0 = r0 * Si[0] + r1 * Si[2];
  const double row1 = r0 * Si[1] + r1 * Si[3];
  return row0 * r0 + row1 * r1;
}

// A measurement of dimension n (1 or 2) on the bound parameters a (and b)
// with values v and the n x n row-major covariance V.
struct MeasDev {
  int n, a, b;
  double v[2];
  double V[4];
};

// The measurement m of a batch (2D on loc0, loc1 without subspace arrays);
// false if its subspace is not one this fitter handles.
__device__ inline bool loadMeas(const DevBatch& in, int64_t m, MeasDev& M) {
  M.n = in.measDims ? in.measDims[m] : 2;
  M.a = in.measIdx ? in.measIdx[2 * m] : 0;
  M.b = in.measIdx ? in.measIdx[2 * m + 1] : 1;
  M.v[0] = in.measVal[2 * m];
  M.v[1] = in.measVal[2 * m + 1];
#pragma unroll
  for (int i = 0; i < 4; ++i) M.V[i] = in.measCov[4 * m + i];
  if (M.n == 1) M.b = 0;
  return (M.n == 1 || M.n == 2) && M.a >= 0 && M.a < 6 &&
         (M.n == 1 || (M.b >= 0 && M.b < 6 && M.b != M.a));
}

// The (0, 1) subspace of a 2D measurement: the fast specialised update
__device__ inline bool isLoc01(const MeasDev& M) { return M.n == 2 && M.a == 0 && M.b == 1; }

// Row i of a packed symmetric matrix
__device__ inline void symRow(const double P[21], int i, double r[6]) {
#pragma unroll
  for (int k = 0; k < 6; ++k) r[k] = P[sym(i, k)];
}

// Entry (i, k) of I - K H for a measurement on the parameters a (, b):
// 1 on the diagonal, minus the gain column of the measured parameter
__device__ inline double ikh(int i, int k, const MeasDev& M, double K0, double K1) {
  const bool isA = k == M.a, isB = M.n == 2 && k == M.b;
  return (i == k ? 1.0 : 0.0) - (isA ? K0 : (isB ? K1 : 0.0));
}
__device__ inline bool ikhNonZero(int i, int k, const MeasDev& M) {
  return k == i || k == M.a || (M.n == 2 && k == M.b);
}

// GainMatrixUpdaterImpl.hpp:21-91 for a measurement of dimension N on any
// subspace.  The products with the projector H are the reference's dense
// products with their zero terms left out: each sum adds its non-zero
// terms in the order of the summation index, which gives the dense sum's
// value exactly (adding a zero product leaves a sum unchanged).
template <int N>
__device__ inline int gainUpdateSub(const double x[6], const double P[21], const MeasDev& M,
                                    bool joseph, double f[6], double C[21], double& chi2) {
  // P H^T (column j = column a or b of P), S = H P H^T + V
  double Pa[6], Pb[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double va = 0.0, vb = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      va = (k == M.a) ? P[sym(i, k)] : va;
      vb = (k == M.b) ? P[sym(i, k)] : vb;
    }
    Pa[i] = va;
    Pb[i] = vb;
  }
  double xa = 0.0, xb = 0.0, Saa = 0.0, Sab = 0.0, Sba = 0.0, Sbb = 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    xa = (k == M.a) ? x[k] : xa;
    xb = (k == M.b) ? x[k] : xb;
    Saa = (k == M.a) ? Pa[k] : Saa;
    Sab = (k == M.a) ? Pb[k] : Sab;
    Sba = (k == M.b) ? Pa[k] : Sba;
    Sbb = (k == M.b) ? Pb[k] : Sbb;
  }
  double Si00, Si01 = 0.0, Si10 = 0.0, Si11 = 0.0;
  if constexpr (N == 1) {
    Si00 = 1.0 / (Saa + M.V[0]);
  } else {
    double o[4];
    inverse2(Saa + M.V[0], Sab + M.V[1], Sba + M.V[2], Sbb + M.V[3], o);
    Si00 = o[0];
    Si01 = o[1];
    Si10 = o[2];
    Si11 = o[3];
  }
  // K = (P H^T) S^-1
  double K0[6], K1[6];
  bool nan = false;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    if constexpr (N == 1) {
      K0[i] = Pa[i] * Si00;
      K1[i] = 0.0;
    } else {
      K0[i] = Pa[i] * Si00 + Pb[i] * Si10;
      K1[i] = Pa[i] * Si01 + Pb[i] * Si11;
    }
    nan = nan || isnan(K0[i]) || isnan(K1[i]);
  }
  if (nan) return ACTS_KF_STATUS_UPDATE_FAILED;
  // filtered = x + K (m - H x), normalised
  const double r0 = M.v[0] - xa, r1 = M.v[1] - xb;
#pragma unroll
  for (int i = 0; i < 6; ++i) f[i] = x[i] + (N == 1 ? K0[i] * r0 : K0[i] * r0 + K1[i] * r1);
  normalizePhiTheta(f[2], f[3]);
  // C = (I - K H) P, or the Joseph form (I - K H) P (I - K H)^T + K V K^T
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double TP[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double acc = 0.0;
      bool started = false;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        if (ikhNonZero(i, k, M)) {
          const double term = ikh(i, k, M, K0[i], K1[i]) * P[sym(k, j)];
          acc = started ? acc + term : term;
          started = true;
        }
      }
      TP[j] = acc;
    }
#pragma unroll
    for (int j = i; j < 6; ++j) {
      if (!joseph) {
        C[sym(i, j)] = TP[j];
      } else {
        double acc = 0.0;
        bool started = false;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          if (ikhNonZero(j, k, M)) {
            const double term = TP[k] * ikh(j, k, M, K0[j], K1[j]);
            acc = started ? acc + term : term;
            started = true;
          }
        }
        // (K V) K^T
        double b;
        if constexpr (N == 1) {
          b = (K0[i] * M.V[0]) * K0[j];
        } else {
          const double kv0 = K0[i] * M.V[0] + K1[i] * M.V[2];
          const double kv1 = K0[i] * M.V[1] + K1[i] * M.V[3];
          b = kv0 * K0[j] + kv1 * K1[j];
        }
        C[sym(i, j)] = acc + b;
      }
    }
  }
  // chi2 = res^T ((I - H K) V)^-1 res on the filtered residual
  double fa = 0.0, fb = 0.0, Ka0 = 0.0, Ka1 = 0.0, Kb0 = 0.0, Kb1 = 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    fa = (k == M.a) ? f[k] : fa;
    fb = (k == M.b) ? f[k] : fb;
    Ka0 = (k == M.a) ? K0[k] : Ka0;
    Ka1 = (k == M.a) ? K1[k] : Ka1;
    Kb0 = (k == M.b) ? K0[k] : Kb0;
    Kb1 = (k == M.b) ? K1[k] : Kb1;
  }
  const double e0 = M.v[0] - fa, e1 = M.v[1] - fb;
  if constexpr (N == 1) {
    const double m00 = (1.0 - Ka0) * M.V[0];
    const double mi = 1.0 / m00;
    chi2 = (e0 * mi) * e0;
  } else {
    const double i00 = 1.0 - Ka0, i01 = 0.0 - Ka1, i10 = 0.0 - Kb0, i11 = 1.0 - Kb1;
    double mi[4];
    inverse2(i00 * M.V[0] + i01 * M.V[2], i00 * M.V[1] + i01 * M.V[3],
             i10 * M.V[0] + i11 * M.V[2], i10 * M.V[1] + i11 * M.V[3], mi);
    const double row0 = e0 * mi[0] + e1 * mi[2];
    const double row1 = e0 * mi[1] + e1 * mi[3];
    chi2 = row0 * e0 + row1 * e1;
  }
  return 0;
}

// GainMatrixUpdater::visitMeasurement (GainMatrixUpdater.cpp:19-29): the
// specialised 2D (loc0, loc1) update, or the dense one of the measurement's
// dimension.
__device__ inline int gainUpdateMeas(const double x[6], const double P[21], const MeasDev& M,
                                     bool joseph, double f[6], double C[21], double& chi2) {
  if (isLoc01(M)) return gainUpdate(x, P, M.v, M.V, joseph, f, C, chi2);
  if (M.n == 1) return gainUpdateSub<1>(x, P, M, joseph, f, C, chi2);
  return gainUpdateSub<2>(x, P, M, joseph, f, C, chi2);
}

// calculatePredictedChi2 (TrackHelpers.hpp:676-694) for any measurement
__device__ inline double predictedChi2Meas(const double x[6], const double P[21],
                                           const MeasDev& M) {
  if (isLoc01(M)) return predictedChi2(x, P, M.v, M.V);
  double ra[6], rb[6];
  symRow(P, 0, ra);
  double xa = x[0], xb = x[0], paa = ra[0], pab = ra[0], pba = ra[0], pbb = ra[0];
#pragma unroll
  for (int k = 1; k < 6; ++k) {
    xa = (M.a == k) ? x[k] : xa;
    xb = (M.b == k) ? x[k] : xb;
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    symRow(P, i, rb);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if (M.a == i && M.a == k) paa = rb[k];
      if (M.a == i && M.b == k) pab = rb[k];
      if (M.b == i && M.a == k) pba = rb[k];
      if (M.b == i && M.b == k) pbb = rb[k];
    }
  }
  if (M.n == 1) {
    const double r0 = M.v[0] - xa;
    const double i00 = 1.0 / (paa + M.V[0]);
    return (r0 * i00) * r0;
  }
  const double r0 = M.v[0] - xa, r1 = M.v[1] - xb;
  double Si[4];
  inverse2(paa + M.V[0], pab + M.V[1], pba + M.V[2], pbb + M.V[3], Si);
  const double row0 = r0 * Si[0] + r1 * Si[2];
  const double row1 = r0 * Si[1] + r1 * Si[3];
  return row0 * r0 + row1 * r1;
}

// 6x6 inverse by partial-pivot LU with all indices static (row swaps done
// with selects so that the matrix stays in registers).
__device__ inline bool inverse6(const double* A, double* X) {
  double lu[36];
  int perm[6];
#pragma unroll
  for (int i = 0; i < 36; ++i) lu[i] = A[i];
#pragma unroll
  for (int i = 0; i < 6; ++i) perm[i] = i;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    int piv = k;
    double big = fabs(lu[6 * k + k]);
#pragma unroll
    for (int i = k + 1; i < 6; ++i) {
      const double v = fabs(lu[6 * i + k]);
      if (v > big) {
        big = v;
        piv = i;
      }
    }
    if (big != 0.0) {
#pragma unroll
      for (int i = k + 1; i < 6; ++i) {
        if (piv == i) {
#pragma unroll
          for (int j = 0; j < 6; ++j) {
            const double tmp = lu[6 * k + j];
            lu[6 * k + j] = lu[6 * i + j];
            lu[6 * i + j] = tmp;
          }
          const int tp = perm[k];
          perm[k] = perm[i];
          perm[i] = tp;
        }
      }
#pragma unroll
      for (int i = k + 1; i < 6; ++i) lu[6 * i + k] = lu[6 * i + k] / lu[6 * k + k];
    }
#pragma unroll
    for (int i = k + 1; i < 6; ++i)
#pragma unroll
      for (int j = k + 1; j < 6; ++j)
        lu[6 * i + j] = lu[6 * i + j] - lu[6 * i + k] * lu[6 * k + j];
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    double b[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) b[i] = (perm[i] == c) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
      for (int i = k + 1; i < 6; ++i) b[i] = b[i] - lu[6 * i + k] * b[k];
#pragma unroll
    for (int k = 5; k >= 0; --k) {
      b[k] = b[k] / lu[6 * k + k];
#pragma unroll
      for (int i = 0; i < k; ++i) b[i] = b[i] - lu[6 * i + k] * b[k];
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) X[6 * i + c] = b[i];
  }
  return true;
}

// GainMatrixSmoother::calculate (GainMatrixSmoother.cpp:18-97); covariances
// packed symmetric, the jacobian full.
__device__ inline int smoothStep(const double xf[6], const double Cf[21],
                                 const double xpN[6], const double CpN[21],
                                 const double xsN[6], const double CsN[21],
                                 const double JN[36], double xs[6],
                                 double Cs[21]) {
  double CfJt[36], Cpi[36], G[36];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double acc = Cf[sym(i, 0)] * JN[6 * j];
#pragma unroll
      for (int k = 1; k < 6; ++k) acc = acc + Cf[sym(i, k)] * JN[6 * j + k];
      CfJt[6 * i + j] = acc;
    }
  {
    double Cp[36];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = 0; j < 6; ++j) Cp[6 * i + j] = CpN[sym(i, j)];
    inverse6(Cp, Cpi);
  }
  bool nan = false;
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double acc = CfJt[6 * i] * Cpi[j];
#pragma unroll
      for (int k = 1; k < 6; ++k) acc = acc + CfJt[6 * i + k] * Cpi[6 * k + j];
      G[6 * i + j] = acc;
      nan = nan || isnan(acc);
    }
  if (nan) return ACTS_KF_STATUS_SMOOTH_FAILED;
  double d[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) d[i] = xsN[i] - xpN[i];
  d[2] = radianSym(d[2]);
  d[3] = radianSym(d[3]);
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double acc = G[6 * i] * d[0];
#pragma unroll
    for (int k = 1; k < 6; ++k) acc = acc + G[6 * i + k] * d[k];
    xs[i] = xf[i] + acc;
  }
  normalizePhiTheta(xs[2], xs[3]);
  double D[21], GD[36];
#pragma unroll
  for (int i = 0; i < 21; ++i) D[i] = CsN[i] - CpN[i];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double acc = G[6 * i] * D[sym(0, j)];
#pragma unroll
      for (int k = 1; k < 6; ++k) acc = acc + G[6 * i + k] * D[sym(k, j)];
      GD[6 * i + j] = acc;
    }
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = i; j < 6; ++j) {
      double acc = GD[6 * i] * G[6 * j];
#pragma unroll
      for (int k = 1; k < 6; ++k) acc = acc + GD[6 * i + k] * G[6 * j + k];
      Cs[sym(i, j)] = Cf[sym(i, j)] + acc;
    }
  return 0;
}

// The reference surface: a PerigeeSurface, i.e. a LineSurface along the
// global z axis through the point c (LineSurface.cpp:57-250).
struct LineRef {
  double c[3];
};

// LineSurface::intersect (LineSurface.cpp:136-186) without bounds
__device__ inline int lineIntersect(const LineRef& L, const V3& pos, const V3& dir,
                                    double tol, double& path) {
  const V3 eb = mk(0., 0., 1.);
  const V3 mab = mk(L.c[0] - pos.x, L.c[1] - pos.y, L.c[2] - pos.z);
  const double eaTeb = dot(dir, eb);
  const double denom = 1 - eaTeb * eaTeb;
  path = INFINITY;
  if (fabs(denom) < fabs(tol)) return kUnreachable;
  const double u = (dot(mab, dir) - dot(mab, eb) * eaTeb) / denom;
  path = u;
  return fabs(u) > fabs(tol) ? kReachable : kOnSurface;
}

// ---------------------------------------------------------------------------
// navigation: DirectNavigator.hpp:280-390, SteppingHelper.hpp:40-78
// ---------------------------------------------------------------------------
struct Nav {
  int seqPos;   // DirectNavigator::State::surfaceIndex; layer navigator: the
                // last target taken from the resolved list (-1: none yet,
                // kNavUnresolved: no list)
  int seqLen;   // the track's sequence length (its state slots)
  bool navBreak;
  int tSurf;    // current target surface (-1 none)
  int tIdx;     // intersection index of the target
};
constexpr int kNavUnresolved = -2;

// Where a lane's layer-navigator record lives (DevScratch::nav and navMask
// of its queue slot) and how many surfaces the geometry has.
struct LayerCtx {
  double* rec;
  uint64_t* mask;
  int64_t stride;
  int32_t nSurfaces;
};

// The layers a track has reached (and its start surface): bit s of word
// s / 64 of its mask.
__device__ inline void markReached(const LayerCtx& lc, int s) {
  lc.mask[(s >> 6) * lc.stride] |= uint64_t{1} << (s & 63);
}
__device__ inline void clearReached(const LayerCtx& lc) {
  for (int w = 0; w < (lc.nSurfaces + 63) / 64; ++w) lc.mask[w * lc.stride] = 0;
}

// The direction the navigation intersects along: the track's direction
// times the propagation direction (backward for the reverse filter).
template <bool Rev>
__device__ inline V3 navDir(const Track& T) {
  return Rev ? mk(-T.dir.x, -T.dir.y, -T.dir.z) : T.dir;
}

// updateSingleSurfaceStatus (SteppingHelper.hpp:40-78): the direct
// navigator's targets with BoundaryTolerance::Infinite, the layer
// navigator's (bounded) with BoundaryTolerance::None.
template <bool Rev = false>
__device__ inline int updateSurfaceStatus(Track& T, const DevSurface& s, int idx,
                                          double tol, bool bounded = false) {
  const V3 d = navDir<Rev>(T);
  const Isect x = bounded ? intersectBounded(s, T.pos, d, tol) : intersect(s, T.pos, d, tol);
  const double path = idx ? x.path[1] : x.path[0];
  const int st = idx ? x.status[1] : x.status[0];
  if (st == kOnSurface) {
    T.cNav = kNotSet;
    updateStep(T.cNav, path);
    return kOnSurface;
  }
  if (st != kUnreachable && path > -DBL_MAX && path < DBL_MAX + kOnSurfaceTolerance) {
    T.cNav = kNotSet;
    updateStep(T.cNav, path);
    return kReachable;
  }
  return kUnreachable;
}

// nextSurface / endOfSurfaces (DirectNavigator.hpp:107-124): the sequence
// in the propagation direction
template <bool Rev = false, typename SeqFn>
__device__ inline void nextTarget(Nav& nv, const Track& T, const DevSurface* geo,
                                  const DevOpts& o, SeqFn seqAt) {
  nv.tSurf = -1;
  if (nv.navBreak) return;
  const int inc = Rev ? -1 : 1;
  nv.seqPos += inc;
  const V3 d = navDir<Rev>(T);
  while (nv.seqPos >= 0 && nv.seqPos < nv.seqLen) {
    const int sid = seqAt(nv.seqPos);
    const Isect x = intersect(geo[sid], T.pos, d, o.surfaceTolerance);
    for (int i = 0; i < x.size; ++i) {
      if (x.status[i] != kUnreachable && x.path[i] > o.nearLimit &&
          x.path[i] < DBL_MAX + kOnSurfaceTolerance) {
        nv.tSurf = sid;
        nv.tIdx = i;
        return;
      }
    }
    nv.seqPos += inc;
  }
  nv.navBreak = true;
}

// A layer of the layer navigator and its approach intersection
// (Layer::resolve and Layer::surfaceOnApproach without an approach
// descriptor, Layer.hpp:147-168, Layer.cpp:224-266): a surface that is
// sensitive or carries material (resolveSensitive, resolveMaterial on,
// resolvePassive off), its first intersection along the line that lies
// inside its bounds, beyond the near limit (the on-surface tolerance,
// NavigatorOptions.hpp:42) and before the far limit (checkPathLength,
// Intersection.hpp:327-337).
__device__ inline bool layerApproach(const DevSurface& s, const V3& pos, const V3& dir,
                                     double& path, int& idx) {
  if (!(s.flags & ACTS_KF_SURFACE_SENSITIVE) && s.mat < 0) return false;
  const Isect x = intersectBounded(s, pos, dir, kOnSurfaceTolerance);
  for (int i = 0; i < x.size; ++i) {
    if (x.status[i] != kUnreachable && x.path[i] > kOnSurfaceTolerance &&
        x.path[i] < DBL_MAX + kOnSurfaceTolerance) {
      path = x.path[i];
      idx = i;
      return true;
    }
  }
  return false;
}

// The layer navigator's next target (Navigator::nextTarget with Gen1
// targeting, Navigator.cpp:251-316,479-540; resolveLayers :826-849;
// TrackingVolume::compatibleLayers, TrackingVolume.cpp:443-487).  After the
// track has reached the layer `on` (and at the start) the layers are
// resolved again from here, as Gen1 does after a volume switch
// (resetAfterVolumeSwitch, Navigator.hpp:245-258); otherwise the next layer
// of the current list is the target.  As in a Gen1 list, which holds each
// layer of a volume once, a layer already reached (or the start surface) is
// not resolved again: where a cylinder and a disc cross each other the track
// would otherwise go back and forth between the two.
// The list of layers resolved from one position and direction, in
// increasing path length (ties by surface index), is not stored: the record
// keeps the position and direction it was resolved from and the path of the
// last target, and the next entry is the layer with the smallest (path,
// index) after the last one, recomputed from the same origin.  A list that
// is used up is resolved again from the current position (renavigation);
// an empty list then ends the navigation.
template <bool Rev = false>
__device__ inline void layerNextTarget(Nav& nv, const Track& T, const DevSurface* geo,
                                       const LayerCtx& lc, int on) {
  nv.tSurf = -1;
  if (nv.navBreak) return;
  if (on >= 0) {
    markReached(lc, on);
    nv.seqPos = kNavUnresolved;
  }
  double* rec = lc.rec;
  const int64_t st = lc.stride;
  for (int attempt = 0; attempt < 2; ++attempt) {
    if (nv.seqPos == kNavUnresolved) {
      const V3 d = navDir<Rev>(T);
      rec[0] = T.pos.x;
      rec[st] = T.pos.y;
      rec[2 * st] = T.pos.z;
      rec[3 * st] = d.x;
      rec[4 * st] = d.y;
      rec[5 * st] = d.z;
      rec[6 * st] = -INFINITY;
      nv.seqPos = -1;
    }
    const V3 o = mk(rec[0], rec[st], rec[2 * st]);
    const V3 d = mk(rec[3 * st], rec[4 * st], rec[5 * st]);
    const double lastPath = rec[6 * st];
    const int lastSurf = nv.seqPos;
    double best = INFINITY;
    int bestS = -1, bestI = 0;
    uint64_t word = 0;
    for (int s = 0; s < lc.nSurfaces; ++s) {
      if ((s & 63) == 0) word = lc.mask[(s >> 6) * st];
      if ((word >> (s & 63)) & 1) continue;
      double p;
      int i;
      if (!layerApproach(geo[s], o, d, p, i)) continue;
      const bool after = p > lastPath || (p == lastPath && s > lastSurf);
      if (after && p < best) {
        best = p;
        bestS = s;
        bestI = i;
      }
    }
    if (bestS >= 0) {
      nv.tSurf = bestS;
      nv.tIdx = bestI;
      nv.seqPos = bestS;
      rec[6 * st] = best;
      return;
    }
    if (attempt == 0) nv.seqPos = kNavUnresolved;
  }
  nv.navBreak = true;
}

// Propagator.ipp:49-79.  `on`: the surface the track has just reached
// (-1 none), which the layer navigator leaves out when it resolves again.
template <bool Rev = false, typename SeqFn>
__device__ inline int getNextTarget(Nav& nv, Track& T, const DevSurface* geo,
                                    const DevOpts& o, SeqFn seqAt, const LayerCtx& lc,
                                    int on) {
  const bool layers = o.navigator != ACTS_KF_NAVIGATOR_DIRECT;
  for (uint32_t i = 0; i < o.maxTargetSkipping; ++i) {
    if (layers)
      layerNextTarget<Rev>(nv, T, geo, lc, on);
    else
      nextTarget<Rev>(nv, T, geo, o, seqAt);
    // Navigator::nextTarget resets the current surface
    on = -1;
    if (nv.tSurf < 0) return 0;
    const int st =
        updateSurfaceStatus<Rev>(T, geo[nv.tSurf], nv.tIdx, o.surfaceTolerance, layers);
    if (st == kOnSurface) continue;
    if (st == kReachable) return 0;
  }
  nv.tSurf = -1;
  return ACTS_KF_STATUS_NEXT_TARGET_LIMIT;
}

// LoopProtection.hpp:19-67 (forward)
template <bool Mixed = false>
__device__ inline void loopProtection(const Track& T, const DevField& F,
                                      const DevOpts& o, double& limit,
                                      bool release) {
  if (!o.loopProtection) return;
  if (release) limit = DBL_MAX;
  // a failed lookup leaves the loop protection off (LoopProtection.hpp)
  V3 B;
  if constexpr (Mixed) {
    F3 b;
    if (fieldAtF(F, tof(T.pos), b)) return;
    B = mk(b.x, b.y, b.z);
  } else {
    if (fieldAt(F, T.pos, B)) return;
  }
  const double Bn = sqrt(dot(B, B));
  if (Bn == 0) return;
  const double p = copysign(static_cast<double>(o.absQ), T.qop) / T.qop;
  const double helixPath = 2 * kPi * p / Bn;
  const double loopLimit = o.loopFraction * helixPath;
  if (fabs(loopLimit) < fabs(limit)) limit = loopLimit;
}

// PathLimitReached::checkAbort (StandardAborters.hpp:41-61)
__device__ inline bool pathAbort(Track& T, double limit, double tol) {
  const double distance = fabs(limit) - fabs(T.pathAcc);
  if (fabs(distance) < fabs(tol)) return true;
  updateStep(T.cActor, distance);
  return false;
}

}  // namespace kf
