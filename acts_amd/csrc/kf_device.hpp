// kf_device.hpp -- the per-track Kalman fit as device code: one lane fits
// one track (forward filter through the direct-navigator surface sequence,
// then the RTS smoother), with all small matrices in registers.
//
// Reference path: KalmanFitter::fit (KalmanFitter.hpp:693-984) driving
// Propagator<EigenStepper<>, DirectNavigator>::propagate (Propagator.ipp:25-210).
#pragma once

#include "kf_common.hpp"

namespace kf {

// Options as the kernels use them (acts_kf_options, resolved on the host).
struct DevOpts {
  double stepTolerance, stepSizeCutOff, initialStepSize, maxStepSize;
  double pathLimit, loopFraction, surfaceTolerance, nearLimit;
  double chi2Cut;
  uint32_t maxTrials, maxSteps, maxTargetSkipping;
  int32_t loopProtection, multipleScattering, energyLoss, useJoseph;
  int32_t absPdg;
  float mass, absQ;
};

// Batch pointers on the device.
struct DevBatch {
  int64_t n;
  const double* startParams;
  const double* startCov;
  const double* startPlane;
  const int64_t* seqOff;
  const int32_t* seqSurf;
  const int64_t* measOff;
  const int32_t* measSurf;
  const double* measVal;
  const double* measCov;
};

// Scratch of the forward pass (states of every track, read back by the
// smoother), tiled per wavefront; see sd()/si() for the layout.
// Covariances are stored as packed upper triangles (kf::sym).
constexpr int kPred = 0, kPCov = 6, kFilt = 27, kFCov = 33, kJac = 54,
              kPath = 90, kChi2 = 91, kScratchComps = 92;
constexpr int kISurf = 0, kIFlags = 1, kIScratchComps = 2;

struct DevScratch {
  double* d;
  int32_t* i;
  int32_t* track;  // per track: [0] status, [1] tip state index
  int64_t stride;  // >= n, multiple of 64
  int32_t maxStates;
};

// ---------------------------------------------------------------------------
// surfaces: CylinderSurface.cpp, DiscSurface.cpp, PlaneSurface.cpp
// ---------------------------------------------------------------------------
enum : int { kUnreachable = 0, kReachable = 1, kOnSurface = 2 };

struct Isect {
  double path[2];
  int status[2];
  int size;
};

KF_HD Isect intersect(const DevSurface& s, const V3& pos,
                                  const V3& dir, double tol) {
  Isect r;
  r.path[0] = r.path[1] = INFINITY;
  r.status[0] = r.status[1] = kUnreachable;
  if (s.type == ACTS_KF_SURFACE_CYLINDER) {
    // CylinderSurface.cpp:198-283, RealQuadraticEquation.hpp:57-66
    r.size = 2;
    const V3 caxis = col(s.R, 2);
    const V3 pc = mk(pos.x - s.t[0], pos.y - s.t[1], pos.z - s.t[2]);
    const V3 pcXcd = cross(pc, caxis);
    const V3 ldXcd = cross(dir, caxis);
    const double a = dot(ldXcd, ldXcd);
    const double b = 2. * dot(ldXcd, pcXcd);
    const double c = dot(pcXcd, pcXcd) - (s.b0 * s.b0);
    const double disc = b * b - 4 * a * c;
    if (disc >= 0) {
      const double sq = sqrt(disc);
      const double q = -0.5 * (b + (b > 0 ? sq : -sq));
      const double first = q / a;
      const double second = c / q;
      const int st1 = fabs(first) < fabs(tol) ? kOnSurface : kReachable;
      if (disc == 0) {
        r.path[0] = r.path[1] = first;
        r.status[0] = r.status[1] = st1;
      } else {
        const int st2 = fabs(second) < fabs(tol) ? kOnSurface : kReachable;
        if (first <= second) {
          r.path[0] = first;
          r.status[0] = st1;
          r.path[1] = second;
          r.status[1] = st2;
        } else {
          r.path[0] = second;
          r.status[0] = st2;
          r.path[1] = first;
          r.status[1] = st1;
        }
      }
    }
    return r;
  }
  // PlanarHelper.hpp:56-77
  r.size = 1;
  const V3 n = col(s.R, 2);
  const double denom = dot(dir, n);
  if (denom == 0) return r;
  const double path =
      dot(n, mk(s.t[0] - pos.x, s.t[1] - pos.y, s.t[2] - pos.z)) / denom;
  r.path[0] = path;
  r.status[0] = fabs(path) < fabs(tol) ? kOnSurface : kReachable;
  return r;
}

// Surface::globalToLocal with the default on-surface tolerance
KF_HD bool globalToLocal(const DevSurface& s, const V3& pos,
                                     double& l0, double& l1) {
  const V3 loc = toLocal(s, pos);
  if (s.type == ACTS_KF_SURFACE_CYLINDER) {
    // CylinderSurface.cpp:125-145
    double inttol = s.b0 * 0.0001;
    if (inttol < 0.01) inttol = 0.01;
    if (fabs(perp(loc) - s.b0) > inttol) return false;
    l0 = s.b0 * atan2(loc.y, loc.x);
    l1 = loc.z;
    return true;
  }
  if (fabs(loc.z) > kOnSurfaceTolerance) return false;
  if (s.type == ACTS_KF_SURFACE_DISC) {
    l0 = perp(loc);
    l1 = atan2(loc.y, loc.x);
  } else {
    l0 = loc.x;
    l1 = loc.y;
  }
  return true;
}

KF_HD V3 localToGlobal(const DevSurface& s, double l0, double l1) {
  if (s.type == ACTS_KF_SURFACE_CYLINDER) {
    const double r = s.b0;
    const double phi = l0 / r;
    return toGlobal(s.R, s.t, mk(r * cos(phi), r * sin(phi), l1));
  }
  if (s.type == ACTS_KF_SURFACE_DISC)
    return toGlobal2(s.R, s.t, l0 * cos(l1), l0 * sin(l1));
  return toGlobal2(s.R, s.t, l0, l1);
}

// CylinderSurface::normal(gctx, position)
KF_HD V3 cylinderNormal(const DevSurface& s, const V3& pos) {
  V3 p3 = toLocal(s, pos);
  p3.z = 0.;
  return rotate(s.R, normalized(p3));
}

// The position part of the reference frame, columns x (fx), y (fy), z (fz).
// Cylinder: CylinderSurface.cpp:93-110; plane/disc: the placement rotation.
struct Frame {
  V3 fx, fy, fz;
};
KF_HD Frame referenceFrame(const DevSurface& s, const V3& pos) {
  Frame f;
  if (s.type == ACTS_KF_SURFACE_CYLINDER) {
    f.fy = col(s.R, 2);
    f.fz = cylinderNormal(s, pos);
    f.fx = normalized(cross(f.fy, f.fz));
  } else {
    f.fx = col(s.R, 0);
    f.fy = col(s.R, 1);
    f.fz = col(s.R, 2);
  }
  return f;
}

KF_HD double pathCorrection(const DevSurface& s, const V3& pos,
                                        const V3& dir) {
  if (s.type == ACTS_KF_SURFACE_CYLINDER) {
    const V3 n = cylinderNormal(s, pos);
    return fabs(1. / dot(n, dir));
  }
  return 1. / fabs(dot(col(s.R, 2), dir));
}

// Local-position columns of boundToFreeJacobian (rows 0..2, cols 0..1) and
// local-position rows of freeToBoundJacobian (rows 0..1, cols 0..2).
// Surface.cpp:249-290, DiscSurface.cpp:197-264.
KF_HD void positionJacobians(const DevSurface& s, const V3& pos,
                                         double b2f[6], double f2b[6]) {
  if (s.type == ACTS_KF_SURFACE_DISC) {
    const V3 posLoc = toLocal(s, pos);
    const double lr = perp(posLoc);
    const double lphi = atan2(posLoc.y, posLoc.x);
    const double lc = cos(lphi), ls = sin(lphi);
    for (int i = 0; i < 3; ++i) {
      const double lx = s.R[3 * i + 0], ly = s.R[3 * i + 1];
      const double c0 = lc * lx + ls * ly;
      const double c1 = lc * ly - ls * lx;
      if (b2f) {
        b2f[2 * i + 0] = c0;
        b2f[2 * i + 1] = lr * c1;
      }
      if (f2b) {
        f2b[i] = c0;
        f2b[3 + i] = c1 / lr;
      }
    }
    return;
  }
  const Frame f = referenceFrame(s, pos);
  if (b2f) {
    b2f[0] = f.fx.x; b2f[1] = f.fy.x;
    b2f[2] = f.fx.y; b2f[3] = f.fy.y;
    b2f[4] = f.fx.z; b2f[5] = f.fy.z;
  }
  if (f2b) {
    f2b[0] = f.fx.x; f2b[1] = f.fx.y; f2b[2] = f.fx.z;
    f2b[3] = f.fy.x; f2b[4] = f.fy.y; f2b[5] = f.fy.z;
  }
}

// direction part of boundToFreeJacobian (JacobianHelpers.hpp:26-41):
// d[0]=(4,2) d[1]=(4,3) d[2]=(5,2) d[3]=(5,3) d[4]=(6,3); (6,2) is 0
KF_HD void directionB2F(const V3& dir, double d[5]) {
  const double cosTheta = dir.z;
  const double sinTheta = cathetus(1, dir.z);
  const double inv = 1. / sinTheta;
  const double cosPhi = dir.x * inv, sinPhi = dir.y * inv;
  d[0] = -dir.y;
  d[1] = cosTheta * cosPhi;
  d[2] = dir.x;
  d[3] = cosTheta * sinPhi;
  d[4] = -sinTheta;
}

// ---------------------------------------------------------------------------
// per-track propagation state (EigenStepper::State, EigenStepper.hpp:81-153)
// ---------------------------------------------------------------------------
struct Track {
  V3 pos;
  double time;
  V3 dir;
  double qop;
  double cov[21];  // packed symmetric (kf::sym)
  // jacToGlobal: position block rows 0..2 x cols 0..1, direction block
  double jp[6];
  double jd[5];
  // jacTransport: top-right rows 0..2 (cols 4..7), (3,7), bottom-right rows 4..6
  double tr[12];
  double tr37;
  double br[12];
  // derivative: (0..2) direction, (3) dt/ds, (4..6) k4
  double der[7];
  double pathAcc;
  int nSteps, nTrials;
  // ConstrainedStep (ConstrainedStep.hpp:44-173)
  double cNav, cActor, cUser, cAcc;
};

__device__ inline double stepValue(const Track& T) {
  const double mn = fmin(fmin(T.cNav, T.cActor), T.cUser);
  const double r = fmin(fabs(mn), T.cAcc);
  return signbit(mn) ? -r : r;
}
__device__ inline void updateStep(double& c, double v) {
  if (fabs(v) < fabs(c)) c = v;
}

__device__ inline void resetTransport(Track& T) {
#pragma unroll
  for (int i = 0; i < 12; ++i) T.tr[i] = 0.0;
#pragma unroll
  for (int i = 0; i < 12; ++i) T.br[i] = 0.0;
  T.br[0] = T.br[5] = T.br[10] = 1.0;
  T.tr37 = 0.0;
#pragma unroll
  for (int i = 0; i < 7; ++i) T.der[i] = 0.0;
}

// InterpolatedBFieldMap::getField (InterpolatedBFieldMap.hpp:335-366,
// Grid.hpp:470-494, Interpolation.hpp:96-120): the cell of the point's bin,
// the last axis collapsed first with (1 - f) * a + f * b.  Returns 0 or
// FIELD_OUT_OF_BOUNDS.
KF_HD int gridField(const DevField& F, const V3& p, V3& out) {
  const double x[3] = {p.x, p.y, p.z};
  int lo[3];
  double f[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (!(x[a] >= F.min[a]) || x[a] >= F.upper[a]) return ACTS_KF_STATUS_FIELD_OUT_OF_BOUNDS;
    // 1 <= bin <= nb: x >= min, and x < upper up to one rounding
    int bin = static_cast<int>(floor((x[a] - F.min[a]) / F.width[a]) + 1);
    bin = bin < 1 ? 1 : (bin > F.nb[a] ? F.nb[a] : bin);
    lo[a] = bin - 1;
    const double ll = F.min[a] + static_cast<double>(bin - 1) * F.width[a];
    const double ur = F.min[a] + static_cast<double>(bin) * F.width[a];
    f[a] = (x[a] - ll) / (ur - ll);
  }
  const int64_t sy = F.nb[2] + 1;
  const int64_t sx = static_cast<int64_t>(F.nb[1] + 1) * (F.nb[2] + 1);
  const double* base = F.v + 3 * (lo[0] * sx + lo[1] * sy + lo[2]);
  // corners (x, y) in canonical order, each with its z pair contiguous
  double c[4][3];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double* q = base + 3 * (((i >> 1) & 1) * sx + (i & 1) * sy);
#pragma unroll
    for (int k = 0; k < 3; ++k) c[i][k] = (1 - f[2]) * q[k] + f[2] * q[3 + k];
  }
  double d[2][3];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) d[i][k] = (1 - f[1]) * c[2 * i][k] + f[1] * c[2 * i + 1][k];
  out = mk((1 - f[0]) * d[0][0] + f[0] * d[1][0], (1 - f[0]) * d[0][1] + f[0] * d[1][1],
           (1 - f[0]) * d[0][2] + f[0] * d[1][2]);
  return 0;
}

KF_HD int fieldAt(const DevField& F, const V3& p, V3& out) {
  if (F.kind == ACTS_KF_FIELD_CONSTANT) {
    out = mk(F.B[0], F.B[1], F.B[2]);
    return 0;
  }
  return gridField(F, p, out);
}

// EigenStepperDefaultExtension.hpp:40-58
__device__ inline V3 kEval(double qop, const V3& dir, double h, const V3& kp,
                           const V3& B) {
  const V3 c = cross(mk(dir.x + h * kp.x, dir.y + h * kp.y, dir.z + h * kp.z), B);
  return mk(qop * c.x, qop * c.y, qop * c.z);
}

__device__ inline double stepScaling(double tol, double err) {
  double x = tol / err;
  x = sqrt(sqrt(x));
  return fmin(fmax(x, 0.25), 4.0);
}

// EigenStepper::step (EigenStepper.ipp:169-388) + transportMatrix
// (EigenStepperDefaultExtension.hpp:137-228), propagation direction +1.
__device__ inline int rkStep(Track& T, const DevField& F, const DevOpts& o) {
  const V3 pos = T.pos;
  const V3 dir = T.dir;
  const double qop = T.qop;
  V3 B1;
  if (const int fs = fieldAt(F, pos, B1)) return fs;
  const V3 c1 = cross(dir, B1);
  const V3 k1 = mk(qop * c1.x, qop * c1.y, qop * c1.z);

  double h = stepValue(T);
  const double initialH = h;
  double h2, half_h, err;
  V3 B2, B3, k2, k3, k4;
  int nTrials = 0;
  while (true) {
    ++nTrials;
    h2 = h * h;
    half_h = h * 0.5;
    const double h28 = h2 * 0.125;
    const V3 p1 = mk((pos.x + half_h * dir.x) + h28 * k1.x,
                     (pos.y + half_h * dir.y) + h28 * k1.y,
                     (pos.z + half_h * dir.z) + h28 * k1.z);
    if (const int fs = fieldAt(F, p1, B2)) return fs;
    k2 = kEval(qop, dir, half_h, k1, B2);
    k3 = kEval(qop, dir, half_h, k2, B2);
    const double h22 = h2 * 0.5;
    const V3 p2 = mk((pos.x + h * dir.x) + h22 * k3.x,
                     (pos.y + h * dir.y) + h22 * k3.y,
                     (pos.z + h * dir.z) + h22 * k3.z);
    if (const int fs = fieldAt(F, p2, B3)) return fs;
    k4 = kEval(qop, dir, h, k3, B3);
    const double e0 = ((k1.x - k2.x) - k3.x) + k4.x;
    const double e1 = ((k1.y - k2.y) - k3.y) + k4.y;
    const double e2 = ((k1.z - k2.z) - k3.z) + k4.z;
    err = h2 * ((fabs(e0) + fabs(e1)) + fabs(e2));
    err = fmax(1e-20, err);
    if (err <= 4.0 * o.stepTolerance) break;
    h *= stepScaling(o.stepTolerance, err);
    if (fabs(h) < fabs(o.stepSizeCutOff)) return ACTS_KF_STATUS_STEP_SIZE_STALLED;
    if (static_cast<uint32_t>(nTrials) > o.maxTrials)
      return ACTS_KF_STATUS_STEP_ADJUSTMENT_FAILED;
  }

  // propagateTime
  const float m = o.mass;
  // ParticleHypothesis::extractMomentum: float(copysign(|q|, q/p)) / (q/p)
  const double p = copysign(static_cast<double>(o.absQ), qop) / qop;
  const double mp = m / p;
  const double dtds = sqrt(1 + mp * mp);
  T.time += h * dtds;
  T.der[3] = dtds;

  // transport matrix D (only the non-trivial blocks)
  const double h6 = h / 6.;
  const double hh6 = (h * h) / 6.;
  const double qh2 = qop * half_h;
  const double qh = qop * h;
  V3 dk1dL = cross(dir, B1);
  V3 dk2dL, dk3dL, dk4dL;
  {
    const V3 a = cross(mk(dir.x + half_h * k1.x, dir.y + half_h * k1.y,
                          dir.z + half_h * k1.z), B2);
    const V3 b = cross(dk1dL, B2);
    dk2dL = mk(a.x + qh2 * b.x, a.y + qh2 * b.y, a.z + qh2 * b.z);
  }
  {
    const V3 a = cross(mk(dir.x + half_h * k2.x, dir.y + half_h * k2.y,
                          dir.z + half_h * k2.z), B2);
    const V3 b = cross(dk2dL, B2);
    dk3dL = mk(a.x + qh2 * b.x, a.y + qh2 * b.y, a.z + qh2 * b.z);
  }
  {
    const V3 a = cross(mk(dir.x + h * k3.x, dir.y + h * k3.y, dir.z + h * k3.z), B3);
    const V3 b = cross(dk3dL, B3);
    dk4dL = mk(a.x + qh * b.x, a.y + qh * b.y, a.z + qh * b.z);
  }
  double dk1dT[9], dk2dT[9], dk3dT[9], dk4dT[9];
  dk1dT[0] = qop * 0.0;     dk1dT[1] = qop * B1.z;  dk1dT[2] = qop * -B1.y;
  dk1dT[3] = qop * -B1.z;   dk1dT[4] = qop * 0.0;   dk1dT[5] = qop * B1.x;
  dk1dT[6] = qop * B1.y;    dk1dT[7] = qop * -B1.x; dk1dT[8] = qop * 0.0;
  auto stage = [&](const double* prev, double hh, const V3& B, double* out) {
    double m3[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) m3[3 * i + j] = (i == j ? 1.0 : 0.0) + hh * prev[3 * i + j];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const V3 c = cross(mk(m3[j], m3[3 + j], m3[6 + j]), B);
      out[j] = qop * c.x;
      out[3 + j] = qop * c.y;
      out[6 + j] = qop * c.z;
    }
  };
  stage(dk1dT, half_h, B2, dk2dT);
  stage(dk2dT, half_h, B2, dk3dT);
  stage(dk3dT, h, B3, dk4dT);

  double dF[12], dG[12];  // rows 0..2 of D_TR and D_BR (cols 0..3)
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int e = 3 * i + j;
      const double id = (i == j ? 1.0 : 0.0);
      dF[4 * i + j] = (id + h6 * ((dk1dT[e] + dk2dT[e]) + dk3dT[e])) * h;
      dG[4 * i + j] = id + h6 * ((dk1dT[e] + 2. * (dk2dT[e] + dk3dT[e])) + dk4dT[e]);
    }
  }
  dF[3] = hh6 * ((dk1dL.x + dk2dL.x) + dk3dL.x);
  dF[7] = hh6 * ((dk1dL.y + dk2dL.y) + dk3dL.y);
  dF[11] = hh6 * ((dk1dL.z + dk2dL.z) + dk3dL.z);
  dG[3] = h6 * ((dk1dL.x + 2. * (dk2dL.x + dk3dL.x)) + dk4dL.x);
  dG[7] = h6 * ((dk1dL.y + 2. * (dk2dL.y + dk3dL.y)) + dk4dL.y);
  dG[11] = h6 * ((dk1dL.z + 2. * (dk2dL.z + dk3dL.z)) + dk4dL.z);
  const double D37 = h * m * m / (p * p * qop * dtds);

  // EigenStepper.ipp:341-347: TR += D_TR * BR ; BR = D_BR * BR
  // (BR row 3 is (0,0,0,1), D_TR row 3 is (0,0,0,D37))
  double ntr[12], nbr[12];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double pt = (dF[4 * i + 0] * T.br[j] + dF[4 * i + 1] * T.br[4 + j]) +
                  dF[4 * i + 2] * T.br[8 + j];
      double pb = (dG[4 * i + 0] * T.br[j] + dG[4 * i + 1] * T.br[4 + j]) +
                  dG[4 * i + 2] * T.br[8 + j];
      if (j == 3) {
        pt = pt + dF[4 * i + 3];
        pb = pb + dG[4 * i + 3];
      }
      ntr[4 * i + j] = T.tr[4 * i + j] + pt;
      nbr[4 * i + j] = pb;
    }
  }
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    T.tr[i] = ntr[i];
    T.br[i] = nbr[i];
  }
  T.tr37 = T.tr37 + D37;

  // EigenStepper.ipp:355-359
  const double h26 = h2 / 6.;
  T.pos = mk(pos.x + (h * dir.x + h26 * ((k1.x + k2.x) + k3.x)),
             pos.y + (h * dir.y + h26 * ((k1.y + k2.y) + k3.y)),
             pos.z + (h * dir.z + h26 * ((k1.z + k2.z) + k3.z)));
  V3 nd = mk(dir.x + h6 * ((k1.x + 2. * (k2.x + k3.x)) + k4.x),
             dir.y + h6 * ((k1.y + 2. * (k2.y + k3.y)) + k4.y),
             dir.z + h6 * ((k1.z + 2. * (k2.z + k3.z)) + k4.z));
  T.dir = normalized(nd);
  T.der[0] = T.dir.x;
  T.der[1] = T.dir.y;
  T.der[2] = T.dir.z;
  T.der[4] = k4.x;
  T.der[5] = k4.y;
  T.der[6] = k4.z;
  T.pathAcc += h;
  ++T.nSteps;
  T.nTrials += nTrials;

  // EigenStepper.ipp:379-385
  const double nextAccuracy = fabs(h * stepScaling(o.stepTolerance, err));
  const double previousAccuracy = fabs(T.cAcc);
  const double initialStepLength = fabs(initialH);
  if (nextAccuracy < initialStepLength || nextAccuracy > previousAccuracy)
    T.cAcc = nextAccuracy;
  return 0;
}

}  // namespace kf
