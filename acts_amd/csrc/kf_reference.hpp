// kf_reference.hpp -- extrapolation of a fitted track to its reference
// surface, after the smoother.
//
// KalmanFitter::fit_impl ends, when KalmanFitterOptions::referenceSurface is
// set, with extrapolateTrackToReferenceSurface (KalmanFitter.hpp:959-971,
// TrackHelpers.hpp:180-383): the first or last measurement state is chosen
// by the strategy from the straight-line distance of its smoothed parameters
// to the surface, and Propagator::propagate(start, target) with the
// ForcedSurfaceReached and PathLimitReached aborters (Propagator.ipp:25-257,
// StandardAborters.hpp:24-164) carries parameters and covariance there, in
// the direction of that distance, with default propagator and stepper
// options and no material.  The reference surface is a PerigeeSurface: a
// LineSurface along the global z axis through `center`
// (LineSurface.cpp:57-250).  A failure fails the fit.
#pragma once

#include "kf_kernel.hpp"

namespace kf {

// LineSurface::referenceFrame (LineSurface.cpp:101-114): columns
// (unitD0, unitZ0, unitDistance)
__device__ inline void lineFrame(const V3& dir, V3& d0, V3& z0, V3& dist) {
  z0 = mk(0., 0., 1.);
  d0 = normalized(cross(z0, dir));
  dist = cross(d0, z0);
}

// LineSurface::globalToLocal (LineSurface.cpp:68-96); the frame is
// orthonormal, its inverse is its transpose
__device__ inline bool lineGlobalToLocal(const LineRef& L, const V3& pos, const V3& dir,
                                         double tol, double& l0, double& l1) {
  V3 d0, z0, dist;
  lineFrame(dir, d0, z0, dist);
  const V3 d = mk(pos.x - L.c[0], pos.y - L.c[1], pos.z - L.c[2]);
  const double x = (d0.x * d.x + d0.y * d.y) + d0.z * d.z;
  const double y = (z0.x * d.x + z0.y * d.y) + z0.z * d.z;
  const double z = (dist.x * d.x + dist.y * d.y) + dist.z * d.z;
  if (fabs(z) > fabs(tol)) return false;
  l0 = x;
  l1 = y;
  return true;
}

// Dense 6x8 / 8x8 / 8x6 helpers in the summation order of the oracle's
// matrix products (k from 0 upwards).
template <int R, int K, int C>
__device__ inline void mulDense(const double* A, const double* B, double* out) {
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int j = 0; j < C; ++j) {
      double acc = A[i * K] * B[j];
#pragma unroll
      for (int k = 1; k < K; ++k) acc = acc + A[i * K + k] * B[k * C + j];
      out[i * C + j] = acc;
    }
}

// boundState on the line with covariance transport
// (CovarianceEngine.cpp:29-165, JacobianEngine.cpp:19-45): the full chain
// F2B * (I + d s) * Jt * B2F, with the line's path derivative, which depends
// on the direction too (LineSurface.cpp:227-250).  Writes the bound
// parameters and the full covariance; returns a status.
// With Sympy, Jt is the identity and B2F is the jacobian M the SympyStepper
// carried.
template <bool Sympy = false>
__device__ inline int lineBoundState(const Track& T, const LineRef& L, double tol,
                                     double params[6], double cov[36]) {
  const V3 pos = T.pos, dir = T.dir;
  double l0 = 0, l1 = 0;
  if (!lineGlobalToLocal(L, pos, dir, tol, l0, l1)) return ACTS_KF_STATUS_NOT_ON_SURFACE;
  params[0] = l0;
  params[1] = l1;
  params[2] = atan2(dir.y, dir.x);
  params[3] = atan2(perp(dir), dir.z);
  params[4] = T.qop;
  params[5] = T.time;

  // freeToPathDerivative of the line
  const V3 pc = mk(pos.x - L.c[0], pos.y - L.c[1], pos.z - L.c[2]);
  const V3 z = mk(0., 0., 1.);
  const double pz = dot(pc, z);
  const double dz = dot(z, dir);
  const double norm = 1 / (1 - dz * dz);
  double s[8] = {norm * (dz * z.x - dir.x), norm * (dz * z.y - dir.y), norm * (dz * z.z - dir.z),
                 0.,
                 norm * (pz * z.x - pc.x),  norm * (pz * z.y - pc.y),  norm * (pz * z.z - pc.z),
                 0.};
  // corr = I + derivative * freeToPath
  double corr[64];
  const double der[8] = {T.der[0], T.der[1], T.der[2], T.der[3],
                         T.der[4], T.der[5], T.der[6], 0.};
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) corr[8 * i + j] = (i == j ? 1.0 : 0.0) + der[i] * s[j];
  // freeToBoundJacobian (Surface.cpp:270-290) with the line's frame
  V3 d0, z0, dist;
  lineFrame(dir, d0, z0, dist);
  const double cosTheta = dir.z;
  const double sinTheta = cathetus(1, dir.z);
  const double inv = 1. / sinTheta;
  const double cosPhi = dir.x * inv, sinPhi = dir.y * inv;
  const double invSinTheta = 1. / sinTheta;
  double F2B[48];
#pragma unroll
  for (int i = 0; i < 48; ++i) F2B[i] = 0.0;
  F2B[0] = d0.x;
  F2B[1] = d0.y;
  F2B[2] = d0.z;
  F2B[8] = z0.x;
  F2B[9] = z0.y;
  F2B[10] = z0.z;
  F2B[2 * 8 + 4] = -sinPhi * invSinTheta;
  F2B[2 * 8 + 5] = cosPhi * invSinTheta;
  F2B[2 * 8 + 6] = 0;
  F2B[3 * 8 + 4] = cosPhi * cosTheta;
  F2B[3 * 8 + 5] = sinPhi * cosTheta;
  F2B[3 * 8 + 6] = -sinTheta;
  F2B[4 * 8 + 7] = 1;
  F2B[5 * 8 + 3] = 1;
  // jacTransport from its blocks
  double Jt[64];
#pragma unroll
  for (int i = 0; i < 64; ++i) Jt[i] = (i % 9 == 0) ? 1.0 : 0.0;
  if constexpr (!Sympy) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        Jt[8 * i + 4 + c] = T.tr[4 * i + c];
        Jt[8 * (4 + i) + 4 + c] = T.br[4 * i + c];
      }
    Jt[8 * 3 + 7] = T.tr37;
  }
  // jacToGlobal from its blocks
  double B2F[48];
#pragma unroll
  for (int i = 0; i < 48; ++i) B2F[i] = 0.0;
  B2F[0] = T.jp[0];
  B2F[1] = T.jp[1];
  B2F[6] = T.jp[2];
  B2F[7] = T.jp[3];
  B2F[12] = T.jp[4];
  B2F[13] = T.jp[5];
  B2F[3 * 6 + 5] = 1;
  B2F[7 * 6 + 4] = 1;
  if constexpr (Sympy) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        B2F[i * 6 + 2 + c] = T.tr[3 * i + c];
        B2F[(4 + i) * 6 + 2 + c] = T.br[3 * i + c];
      }
    B2F[3 * 6 + 4] = T.tr37;
  } else {
    B2F[4 * 6 + 2] = T.jd[0];
    B2F[4 * 6 + 3] = T.jd[1];
    B2F[5 * 6 + 2] = T.jd[2];
    B2F[5 * 6 + 3] = T.jd[3];
    B2F[6 * 6 + 3] = T.jd[4];
  }
  double A[48], B[48], J[36];
  mulDense<6, 8, 8>(F2B, corr, A);
  mulDense<6, 8, 8>(A, Jt, B);
  mulDense<6, 8, 6>(B, B2F, J);
  double C[36], JC[36];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) C[6 * i + j] = T.cov[sym(i, j)];
  mulDense<6, 6, 6>(J, C, JC);
  // (J C) J^T
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double acc = JC[6 * i] * J[6 * j];
#pragma unroll
      for (int k = 1; k < 6; ++k) acc = acc + JC[6 * i + k] * J[6 * j + k];
      cov[6 * i + j] = acc;
    }
  return 0;
}

// EigenStepper::initialize (EigenStepper.ipp:37-70) from bound parameters
// and a packed covariance on surface s
template <bool Sympy = false>
__device__ inline void initializeFromBound(Track& T, const DevSurface& s, const double p[6],
                                           const double c[21], const DevOpts& o) {
  const double sinTheta = sin(p[3]);
  T.dir = mk(cos(p[2]) * sinTheta, sin(p[2]) * sinTheta, cos(p[3]));
  T.pos = localToGlobal(s, p[0], p[1]);
  T.time = p[5];
  T.qop = p[4];
#pragma unroll
  for (int i = 0; i < 21; ++i) T.cov[i] = c[i];
  double b2f[6];
  positionJacobians(s, T.pos, b2f, nullptr);
#pragma unroll
  for (int i = 0; i < 6; ++i) T.jp[i] = b2f[i];
  directionB2F(T.dir, T.jd);
  resetTransport(T);
  if constexpr (Sympy) sympyResetM(T);
  T.pathAcc = 0;
  T.nSteps = 0;
  T.nTrials = 0;
  T.cNav = kNotSet;
  T.cActor = kNotSet;
  T.cUser = o.maxStepSize;
  T.cAcc = o.initialStepSize;
}

// Propagator::propagate(start, target) to the line in direction `direction`
// with ForcedSurfaceReached + PathLimitReached (loop protection on);
// returns a status, T holds the state at the end.
template <bool Mixed, bool Sympy = false>
__device__ inline int propagateToLine(Track& T, const LineRef& L, const DevField& F,
                                      const DevOpts& o, double direction) {
  double limit = o.pathLimit;
  // setupLoopProtection: |limit| is what counts (PathLimitReached uses
  // |limit| - |path|), so the direction's sign does not enter
  loopProtection<Mixed>(T, F, o, limit, false);
  auto checkAbort = [&]() -> bool {
    // ForcedSurfaceReached::checkAbort: the intersection along the
    // propagation direction; near limit lowest, far limit max
    const V3 dd = mk(direction * T.dir.x, direction * T.dir.y, direction * T.dir.z);
    double path;
    const int st = lineIntersect(L, T.pos, dd, o.surfaceTolerance, path);
    if (st != kUnreachable && path > -DBL_MAX && path < DBL_MAX + kOnSurfaceTolerance) {
      T.cActor = kNotSet;
      updateStep(T.cActor, path);
    }
    if (st == kOnSurface) return true;
    return pathAbort(T, limit, o.surfaceTolerance);
  };
  if (checkAbort()) return 0;
  for (uint32_t steps = 0; steps < o.maxSteps; ++steps) {
    int sst;
    if constexpr (Mixed && Sympy)
      sst = rkStepSympyMixed(T, F, o, direction);
    else if constexpr (Mixed)
      sst = rkStepMixed(T, F, o, direction);
    else if constexpr (Sympy)
      sst = rkStepSympy(T, F, o, direction);
    else
      sst = rkStep(T, F, o, direction);
    if (sst) return sst;
    T.cNav = kNotSet;
    T.cActor = kNotSet;
    if (checkAbort()) return 0;
  }
  return ACTS_KF_STATUS_STEP_COUNT_LIMIT;
}

// extrapolateTrackToReferenceSurface for track t (states in the scratch,
// smoothed first/last measurement states in `ref`); returns a status and
// writes the parameters and covariance on the line.
template <bool Mixed, bool Sympy = false>
__device__ inline int extrapolateTrack(int64_t q, const DevScratch& sc, const double* ref,
                                       const DevSurface* geo, const DevField& F,
                                       const DevOpts& o, const LineRef& L, int32_t strategy,
                                       double params[6], double cov[36]) {
  const int first = sc.track[2 * sc.stride + q];
  const int last = sc.track[3 * sc.stride + q];
  if (first < 0) return ACTS_KF_STATUS_NO_COMPATIBLE_STATE;
  const double* rf = ref + kRefComps * q;
  // the straight-line distance from a state's smoothed parameters
  auto distanceFrom = [&](int k, const double* p, double& path) -> bool {
    const DevSurface& s = geo[si(sc, k, kISurf, q)];
    const double sinTheta = sin(p[3]);
    const V3 dir = mk(cos(p[2]) * sinTheta, sin(p[2]) * sinTheta, cos(p[3]));
    const V3 pos = localToGlobal(s, p[0], p[1]);
    return lineIntersect(L, pos, dir, kOnSurfaceTolerance, path) != kUnreachable;
  };
  int k = -1;
  const double* p = nullptr;
  double distance = 0;
  if (strategy == ACTS_KF_EXTRAPOLATE_FIRST || strategy == ACTS_KF_EXTRAPOLATE_LAST) {
    const bool isFirst = strategy == ACTS_KF_EXTRAPOLATE_FIRST;
    k = isFirst ? first : last;
    p = isFirst ? rf : rf + 27;
    if (!distanceFrom(k, p, distance)) return ACTS_KF_STATUS_REFERENCE_UNREACHABLE;
  } else {
    double df, dl;
    const bool vf = distanceFrom(first, rf, df);
    const bool vl = distanceFrom(last, rf + 27, dl);
    const double af = fabs(df), al = fabs(dl);
    if (vf && af <= al) {
      k = first;
      p = rf;
      distance = df;
    } else if (vl && al <= af) {
      k = last;
      p = rf + 27;
      distance = dl;
    } else {
      return ACTS_KF_STATUS_REFERENCE_UNREACHABLE;
    }
  }
  // Direction::fromScalarZeroAsPositive
  const double direction = distance >= 0 ? 1.0 : -1.0;
  Track T;
  initializeFromBound<Sympy>(T, geo[si(sc, k, kISurf, q)], p, p + 6, o);
  if (const int st = propagateToLine<Mixed, Sympy>(T, L, F, o, direction)) return st;
  return lineBoundState<Sympy>(T, L, o.surfaceTolerance, params, cov);
}

}  // namespace kf
