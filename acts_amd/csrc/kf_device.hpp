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
  int32_t navigator;  // ACTS_KF_NAVIGATOR_*
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
  const uint8_t* measDims;  // per measurement (null: all 2D)
  const uint8_t* measIdx;   // 2 per measurement (null: loc0, loc1)
};

// Scratch of the forward pass (states of every track, read back by the
// smoother), tiled per wavefront; see sd()/si() for the layout.
// Covariances are stored as packed upper triangles (kf::sym).
constexpr int kPred = 0, kPCov = 6, kFilt = 27, kFCov = 33, kJac = 54,
              kPath = 90, kChi2 = 91, kScratchComps = 92;
// ints: surface, type flags, index of the state's measurement in the
// track's measurements (-1: none)
constexpr int kISurf = 0, kIFlags = 1, kIMeas = 2, kIScratchComps = 3;
// per track with a reference surface: smoothed parameters (6) and packed
// covariance (21) of the first, then of the last measurement state
constexpr int kRefComps = 54;
// layer navigator record: the position and navigation direction its layer
// list was resolved from, and the path length of the current target
constexpr int kNavComps = 7;

struct DevScratch {
  double* d;
  int32_t* i;
  int32_t* track;  // per track: [0] status, [1] tip state index, [2] [3]
                   // first and last measurement (not outlier) state index
  double* ref;     // per track with a reference surface: the smoothed
                   // first and last measurement states (null: none)
  double* nav;     // layer navigator: per queue slot, component j at
                   // nav[j * stride + q] (null: direct navigator)
  uint64_t* navMask;  // and its reached layers, word w at navMask[w * stride + q]
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

// The bounds check of Surface::intersect with BoundaryTolerance::None at
// the intersection point x: full-azimuth cylinders compare the local z with
// the half length plus the tolerance (CylinderSurface.cpp:242-257), full
// discs the squared distance from the centre with rmin^2 and rmax^2
// (DiscSurface.cpp:286-293), planes the local x and y with the half lengths
// of their RectangleBounds (inclusive; b0 = b1 = 0: no bounds).
KF_HD bool insideNavBounds(const DevSurface& s, const V3& x, double tol) {
  const V3 rel = mk(x.x - s.t[0], x.y - s.t[1], x.z - s.t[2]);
  if (s.type == ACTS_KF_SURFACE_CYLINDER) {
    const double cz = dot(rel, col(s.R, 2));
    return fabs(cz) < fabs(s.b1 + tol);
  }
  if (s.type == ACTS_KF_SURFACE_DISC) {
    const double r2 = dot(rel, rel);
    return r2 >= s.b0 * s.b0 && r2 <= s.b1 * s.b1;
  }
  if (s.type == ACTS_KF_SURFACE_PLANE && s.b0 > 0 && s.b1 > 0) {
    const double lx = dot(rel, col(s.R, 0)), ly = dot(rel, col(s.R, 1));
    return -s.b0 <= lx && lx <= s.b0 && -s.b1 <= ly && ly <= s.b1;
  }
  return true;
}

// Surface::intersect with BoundaryTolerance::None: solutions outside the
// bounds are unreachable.
KF_HD Isect intersectBounded(const DevSurface& s, const V3& pos, const V3& dir, double tol) {
  Isect r = intersect(s, pos, dir, tol);
  for (int i = 0; i < r.size; ++i) {
    if (r.status[i] == kUnreachable) continue;
    const V3 x = mk(pos.x + r.path[i] * dir.x, pos.y + r.path[i] * dir.y,
                    pos.z + r.path[i] * dir.z);
    if (!insideNavBounds(s, x, tol)) r.status[i] = kUnreachable;
  }
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
// The cell of p and its interpolation weights; 0 or FIELD_OUT_OF_BOUNDS.
KF_HD int gridLocate(const DevField& F, const V3& p, int lo[3], double f[3]) {
  const double x[3] = {p.x, p.y, p.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (!(x[a] >= F.min[a]) || x[a] >= F.upper[a]) return ACTS_KF_STATUS_FIELD_OUT_OF_BOUNDS;
    // 1 <= bin <= nb: x >= min, and x < upper up to one rounding
    const double u = F.recip ? divRecip(x[a] - F.min[a], F.width[a], F.rw[a])
                             : (x[a] - F.min[a]) / F.width[a];
    int bin = static_cast<int>(floor(u) + 1);
    bin = bin < 1 ? 1 : (bin > F.nb[a] ? F.nb[a] : bin);
    lo[a] = bin - 1;
    const double ll = F.min[a] + static_cast<double>(bin - 1) * F.width[a];
    if (F.recip) {
      // ur - ll is den[a] in every bin
      f[a] = divRecip(x[a] - ll, F.den[a], F.rden[a]);
    } else {
      const double ur = F.min[a] + static_cast<double>(bin) * F.width[a];
      f[a] = (x[a] - ll) / (ur - ll);
    }
  }
  return 0;
}

KF_HD int gridField(const DevField& F, const V3& p, V3& out) {
  int lo[3];
  double f[3];
  if (const int st = gridLocate(F, p, lo, f)) return st;
  // corners (x, y) in canonical order, each with its z pair: 6 values from
  // the cell record, or from two neighbouring points of the point array
  double c[4][3];
  if (F.cells) {
    const int64_t cell = (static_cast<int64_t>(lo[0]) * F.nb[1] + lo[1]) * F.nb[2] + lo[2];
    double v[kCellValues];
    if (F.cache && F.cacheIdx[0] == cell) {
      // the lane's last cell: the same values without the gather
#pragma unroll
      for (int j = 0; j < kCellValues; ++j) v[j] = F.cache[j * kCellCacheStride];
    } else {
      const double* q = static_cast<const double*>(
          __builtin_assume_aligned(F.cells + kCellDoubles * cell, 64));
#pragma unroll
      for (int j = 0; j < kCellValues; ++j) v[j] = q[j];
      if (F.cache) {
#pragma unroll
        for (int j = 0; j < kCellValues; ++j) F.cache[j * kCellCacheStride] = v[j];
        F.cacheIdx[0] = cell;
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) c[i][k] = (1 - f[2]) * v[6 * i + k] + f[2] * v[6 * i + 3 + k];
  } else {
    const int64_t sy = F.nb[2] + 1;
    const int64_t sx = static_cast<int64_t>(F.nb[1] + 1) * (F.nb[2] + 1);
    const double* base = F.v + 3 * (lo[0] * sx + lo[1] * sy + lo[2]);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double* q = base + 3 * (((i >> 1) & 1) * sx + (i & 1) * sy);
#pragma unroll
      for (int k = 0; k < 3; ++k) c[i][k] = (1 - f[2]) * q[k] + f[2] * q[3 + k];
    }
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

// Everything one accepted RK step leaves for the transport-matrix update
// and the state update.
struct RkStage {
  V3 pos, dir, B1, B2, B3, k1, k2, k3, k4;
  double qop, h, h2, half_h, err, initialH;
  int nTrials;
};

// EigenStepper::step (EigenStepper.ipp:169-388): the adaptive RKN4 trials
// until the error estimate is accepted.  Returns 0 or a status.  `propDir`
// is the propagation direction (+1 or -1): the step constraints are path
// lengths along it, and the step is their value times the direction
// (EigenStepper.ipp:270).
__device__ inline int rkTrials(const Track& T, const DevField& F, const DevOpts& o, RkStage& S,
                               double propDir = 1.0) {
  S.pos = T.pos;
  S.dir = T.dir;
  S.qop = T.qop;
  const V3& pos = S.pos;
  const V3& dir = S.dir;
  const double qop = S.qop;
  if (const int fs = fieldAt(F, pos, S.B1)) return fs;
  const V3 c1 = cross(dir, S.B1);
  S.k1 = mk(qop * c1.x, qop * c1.y, qop * c1.z);
  const V3& k1 = S.k1;

  double h = stepValue(T) * propDir;
  S.initialH = h;
  double h2, half_h, err;
  int nTrials = 0;
  while (true) {
    ++nTrials;
    h2 = h * h;
    half_h = h * 0.5;
    const double h28 = h2 * 0.125;
    const V3 p1 = mk((pos.x + half_h * dir.x) + h28 * k1.x,
                     (pos.y + half_h * dir.y) + h28 * k1.y,
                     (pos.z + half_h * dir.z) + h28 * k1.z);
    if (const int fs = fieldAt(F, p1, S.B2)) return fs;
    S.k2 = kEval(qop, dir, half_h, k1, S.B2);
    S.k3 = kEval(qop, dir, half_h, S.k2, S.B2);
    const double h22 = h2 * 0.5;
    const V3 p2 = mk((pos.x + h * dir.x) + h22 * S.k3.x,
                     (pos.y + h * dir.y) + h22 * S.k3.y,
                     (pos.z + h * dir.z) + h22 * S.k3.z);
    if (const int fs = fieldAt(F, p2, S.B3)) return fs;
    S.k4 = kEval(qop, dir, h, S.k3, S.B3);
    const double e0 = ((k1.x - S.k2.x) - S.k3.x) + S.k4.x;
    const double e1 = ((k1.y - S.k2.y) - S.k3.y) + S.k4.y;
    const double e2 = ((k1.z - S.k2.z) - S.k3.z) + S.k4.z;
    err = h2 * ((fabs(e0) + fabs(e1)) + fabs(e2));
    err = fmax(1e-20, err);
    if (err <= 4.0 * o.stepTolerance) break;
    h *= stepScaling(o.stepTolerance, err);
    if (fabs(h) < fabs(o.stepSizeCutOff)) return ACTS_KF_STATUS_STEP_SIZE_STALLED;
    if (static_cast<uint32_t>(nTrials) > o.maxTrials)
      return ACTS_KF_STATUS_STEP_ADJUSTMENT_FAILED;
  }
  S.h = h;
  S.h2 = h2;
  S.half_h = half_h;
  S.err = err;
  S.nTrials = nTrials;
  return 0;
}

// propagateTime (EigenStepperDefaultExtension.hpp:112-125); returns dt/ds
// and the momentum p used by the transport matrix.
__device__ inline double rkTime(Track& T, const DevOpts& o, const RkStage& S, double& p) {
  const float m = o.mass;
  // ParticleHypothesis::extractMomentum: float(copysign(|q|, q/p)) / (q/p)
  p = copysign(static_cast<double>(o.absQ), S.qop) / S.qop;
  const double mp = m / p;
  const double dtds = sqrt(1 + mp * mp);
  T.time += S.h * dtds;
  T.der[3] = dtds;
  return dtds;
}

// d(k_i)/dT of the four stages, column j (j = 0..2) only.
__device__ inline void dkdTColumn(const RkStage& S, int j, double c1[3], double c2[3],
                                  double c3[3], double c4[3]) {
  const double qop = S.qop;
  const V3& B1 = S.B1;
  // column j of dk1dT = qop * (e_j x B1)
  const double d1[9] = {qop * 0.0, qop * B1.z, qop * -B1.y, qop * -B1.z, qop * 0.0,
                        qop * B1.x, qop * B1.y, qop * -B1.x, qop * 0.0};
  c1[0] = d1[j];
  c1[1] = d1[3 + j];
  c1[2] = d1[6 + j];
  auto stage = [&](const double* prev, double hh, const V3& B, double* out) {
    const double m0 = (j == 0 ? 1.0 : 0.0) + hh * prev[0];
    const double m1 = (j == 1 ? 1.0 : 0.0) + hh * prev[1];
    const double m2 = (j == 2 ? 1.0 : 0.0) + hh * prev[2];
    const V3 c = cross(mk(m0, m1, m2), B);
    out[0] = qop * c.x;
    out[1] = qop * c.y;
    out[2] = qop * c.z;
  };
  stage(c1, S.half_h, S.B2, c2);
  stage(c2, S.half_h, S.B2, c3);
  stage(c3, S.h, S.B3, c4);
}

// d(k_i)/d(q/p) of the four stages.
__device__ inline void dkdL(const RkStage& S, V3& d1, V3& d2, V3& d3, V3& d4) {
  const V3& dir = S.dir;
  const double qh2 = S.qop * S.half_h;
  const double qh = S.qop * S.h;
  d1 = cross(dir, S.B1);
  {
    const V3 a = cross(mk(dir.x + S.half_h * S.k1.x, dir.y + S.half_h * S.k1.y,
                          dir.z + S.half_h * S.k1.z), S.B2);
    const V3 b = cross(d1, S.B2);
    d2 = mk(a.x + qh2 * b.x, a.y + qh2 * b.y, a.z + qh2 * b.z);
  }
  {
    const V3 a = cross(mk(dir.x + S.half_h * S.k2.x, dir.y + S.half_h * S.k2.y,
                          dir.z + S.half_h * S.k2.z), S.B2);
    const V3 b = cross(d2, S.B2);
    d3 = mk(a.x + qh2 * b.x, a.y + qh2 * b.y, a.z + qh2 * b.z);
  }
  {
    const V3 a = cross(mk(dir.x + S.h * S.k3.x, dir.y + S.h * S.k3.y, dir.z + S.h * S.k3.z),
                       S.B3);
    const V3 b = cross(d3, S.B3);
    d4 = mk(a.x + qh * b.x, a.y + qh * b.y, a.z + qh * b.z);
  }
}

// Rows 0..2 of the non-trivial transport blocks D_TR (dF) and D_BR (dG),
// column j (0..2 from dk/dT, 3 from dk/dL): out[i] = dF[4i+j], out[3+i] = dG[4i+j].
__device__ inline void dFdGColumn(const RkStage& S, int j, const double* c1, const double* c2,
                                  const double* c3, const double* c4, double out[6]) {
  const double h = S.h;
  const double h6 = h / 6.;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double id = (i == j ? 1.0 : 0.0);
    out[i] = (id + h6 * ((c1[i] + c2[i]) + c3[i])) * h;
    out[3 + i] = id + h6 * ((c1[i] + 2. * (c2[i] + c3[i])) + c4[i]);
  }
}
__device__ inline void dFdGLambda(const RkStage& S, const V3& d1, const V3& d2, const V3& d3,
                                  const V3& d4, double out[6]) {
  const double h = S.h;
  const double h6 = h / 6.;
  const double hh6 = (h * h) / 6.;
  out[0] = hh6 * ((d1.x + d2.x) + d3.x);
  out[1] = hh6 * ((d1.y + d2.y) + d3.y);
  out[2] = hh6 * ((d1.z + d2.z) + d3.z);
  out[3] = h6 * ((d1.x + 2. * (d2.x + d3.x)) + d4.x);
  out[4] = h6 * ((d1.y + 2. * (d2.y + d3.y)) + d4.y);
  out[5] = h6 * ((d1.z + 2. * (d2.z + d3.z)) + d4.z);
}
__device__ inline double transportD37(const DevOpts& o, const RkStage& S, double p, double dtds) {
  const float m = o.mass;
  return S.h * m * m / (p * p * S.qop * dtds);
}

// EigenStepper.ipp:341-347: TR += D_TR * BR ; BR = D_BR * BR, entry (i, j)
// (BR row 3 is (0,0,0,1), D_TR row 3 is (0,0,0,D37)).  dF/dG rows i hold the
// columns 0..3 of row i.
__device__ inline void blockEntry(int i, int j, const double* dFrow, const double* dGrow,
                                  const double bj[3], double tr, double& ntr, double& nbr) {
  // bj = column j of BR (rows 0..2)
  double pt = (dFrow[0] * bj[0] + dFrow[1] * bj[1]) + dFrow[2] * bj[2];
  double pb = (dGrow[0] * bj[0] + dGrow[1] * bj[1]) + dGrow[2] * bj[2];
  if (j == 3) {
    pt = pt + dFrow[3];
    pb = pb + dGrow[3];
  }
  ntr = tr + pt;
  nbr = pb;
  (void)i;
}

// The transport-matrix update of one accepted step, in one lane.
__device__ inline void rkTransport(Track& T, const DevOpts& o, const RkStage& S, double p,
                                   double dtds) {
  double col[4][6];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double c1[3], c2[3], c3[3], c4[3];
    dkdTColumn(S, j, c1, c2, c3, c4);
    dFdGColumn(S, j, c1, c2, c3, c4, col[j]);
  }
  V3 d1, d2, d3, d4;
  dkdL(S, d1, d2, d3, d4);
  dFdGLambda(S, d1, d2, d3, d4, col[3]);
  // column j of both blocks at a time: the new column depends only on the
  // old column j of BR, so the blocks are updated in place
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const double bj[3] = {T.br[j], T.br[4 + j], T.br[8 + j]};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double dFrow[4] = {col[0][i], col[1][i], col[2][i], col[3][i]};
      const double dGrow[4] = {col[0][3 + i], col[1][3 + i], col[2][3 + i], col[3][3 + i]};
      double ntr, nbr;
      blockEntry(i, j, dFrow, dGrow, bj, T.tr[4 * i + j], ntr, nbr);
      T.tr[4 * i + j] = ntr;
      T.br[4 * i + j] = nbr;
    }
  }
  T.tr37 = T.tr37 + transportD37(o, S, p, dtds);
}

// EigenStepper.ipp:355-385: position, direction, derivative, path,
// statistics, accuracy.
__device__ inline void rkApply(Track& T, const DevOpts& o, const RkStage& S) {
  const V3& pos = S.pos;
  const V3& dir = S.dir;
  const double h = S.h;
  const double h6 = h / 6.;
  const double h26 = S.h2 / 6.;
  T.pos = mk(pos.x + (h * dir.x + h26 * ((S.k1.x + S.k2.x) + S.k3.x)),
             pos.y + (h * dir.y + h26 * ((S.k1.y + S.k2.y) + S.k3.y)),
             pos.z + (h * dir.z + h26 * ((S.k1.z + S.k2.z) + S.k3.z)));
  V3 nd = mk(dir.x + h6 * ((S.k1.x + 2. * (S.k2.x + S.k3.x)) + S.k4.x),
             dir.y + h6 * ((S.k1.y + 2. * (S.k2.y + S.k3.y)) + S.k4.y),
             dir.z + h6 * ((S.k1.z + 2. * (S.k2.z + S.k3.z)) + S.k4.z));
  T.dir = normalized(nd);
  T.der[0] = T.dir.x;
  T.der[1] = T.dir.y;
  T.der[2] = T.dir.z;
  T.der[4] = S.k4.x;
  T.der[5] = S.k4.y;
  T.der[6] = S.k4.z;
  T.pathAcc += h;
  ++T.nSteps;
  T.nTrials += S.nTrials;
  const double nextAccuracy = fabs(h * stepScaling(o.stepTolerance, S.err));
  const double previousAccuracy = fabs(T.cAcc);
  const double initialStepLength = fabs(S.initialH);
  if (nextAccuracy < initialStepLength || nextAccuracy > previousAccuracy)
    T.cAcc = nextAccuracy;
}

// EigenStepper::step (EigenStepper.ipp:169-388) + transportMatrix
// (EigenStepperDefaultExtension.hpp:137-228), propagation direction +1.
// Without `transport` the step leaves the transport blocks (tr, tr37, br)
// alone: the trajectory and the step-size control do not depend on them.
__device__ inline int rkStep(Track& T, const DevField& F, const DevOpts& o,
                             double propDir = 1.0, bool transport = true) {
  RkStage S;
  if (const int st = rkTrials(T, F, o, S, propDir)) return st;
  double p;
  const double dtds = rkTime(T, o, S, p);
  if (transport) rkTransport(T, o, S, p, dtds);
  rkApply(T, o, S);
  return 0;
}

// ---------------------------------------------------------------------------
// SympyStepper (SympyStepper.cpp:141-305, vacuum kernel of
// generate_sympy_stepper.py:157-444).  The stepper carries the
// bound-to-free jacobian M itself instead of an 8x8 transport matrix; M
// lives in the Track's transport fields: tr[3 i + j] = M(i, 2 + j) (position
// rows, phi / theta / q/p columns), br[3 i + j] = M(4 + i, 2 + j) (direction
// rows), tr37 = M(3, 4); its loc columns are jp, M(3, 5) = M(7, 4) = 1.
// ---------------------------------------------------------------------------

// M = the surface's boundToFreeJacobian (jp, jd already set)
__device__ inline void sympyResetM(Track& T) {
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    T.tr[i] = 0.0;
    T.br[i] = 0.0;
  }
  T.br[0] = T.jd[0];
  T.br[1] = T.jd[1];
  T.br[3] = T.jd[2];
  T.br[4] = T.jd[3];
  T.br[7] = T.jd[4];
  T.tr37 = 0.0;
#pragma unroll
  for (int i = 0; i < 7; ++i) T.der[i] = 0.0;
}

__device__ inline V3 add3(const V3& a, const V3& b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline V3 sub3(const V3& a, const V3& b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline V3 scale3(double s, const V3& a) { return mk(s * a.x, s * a.y, s * a.z); }

// One step: trials of the ATLAS-form RK4 over half-step bend vectors
// b_i = (h q/p / 2) B_i until the error estimate 2|h| |(dir + dir2 + kick4)
// - (dir3 + dir4)|_1 (= h^2 |k1 - k2 - k3 + k4|_1) is within 4 x the
// tolerance; then the new position, normalised direction and time, and
// (with `transport`) M's phi, theta and q/p columns through the same
// recursion (tangents of the stages) and the path derivatives.
__device__ inline int rkStepSympy(Track& T, const DevField& F, const DevOpts& o,
                                  double propDir = 1.0, bool transport = true) {
  const V3 pos = T.pos;
  const V3 dir = T.dir;
  const double qop = T.qop;
  const double pabs = copysign(static_cast<double>(o.absQ), qop) / qop;
  const double m = o.mass;
  V3 B1;
  if (const int fs = fieldAt(F, pos, B1)) return fs;
  const double initialH = stepValue(T) * propDir;
  double h = initialH, err = 0, hThird = 0;
  int nTrials = 0;
  V3 bend1, bend2, bend3, kick1, dir2, dir3, dir4, dirEnd, kick4;
  while (true) {
    ++nTrials;
    const double halfHQop = h * qop / 2;
    hThird = h / 3;
    const double hQuarter = h / 4;
    bend1 = scale3(halfHQop, B1);
    kick1 = cross(dir, bend1);
    dir2 = add3(kick1, dir);
    const V3 dirHalfSum = add3(dir2, dir);
    const V3 pos2 = add3(pos, scale3(hQuarter, dirHalfSum));
    V3 B2;
    if (const int fs = fieldAt(F, pos2, B2)) return fs;
    bend2 = scale3(halfHQop, B2);
    dir3 = add3(dir, cross(dir2, bend2));
    dir4 = add3(dir, cross(dir3, bend2));
    dirEnd = sub3(scale3(2, dir4), dir);
    const V3 pos3 = add3(pos, scale3(h, dir4));
    V3 B3;
    if (const int fs = fieldAt(F, pos3, B3)) return fs;
    bend3 = scale3(halfHQop, B3);
    kick4 = cross(dirEnd, bend3);
    const V3 e = sub3(add3(dirHalfSum, kick4), add3(dir3, dir4));
    err = 2 * fabs(h) * ((fabs(e.x) + fabs(e.y)) + fabs(e.z));
    const bool ok = !(err > 4.0 * o.stepTolerance);
    err = fmax(1e-20, err);
    if (ok) break;
    h *= stepScaling(o.stepTolerance, err);
    if (fabs(h) < fabs(o.stepSizeCutOff)) return ACTS_KF_STATUS_STEP_SIZE_STALLED;
    if (static_cast<uint32_t>(nTrials) > o.maxTrials) return ACTS_KF_STATUS_STEP_ADJUSTMENT_FAILED;
  }
  const V3 newPos = add3(pos, scale3(hThird, add3(add3(dir2, dir3), dir4)));
  const V3 newDirX3 = add3(scale3(2, dir3), add3(add3(kick1, dirEnd), kick4));
  const double invNorm = 1 / sqrt(dot(newDirX3, newDirX3));
  const V3 newDir = scale3(invNorm, newDirX3);
  const double dtds = sqrt(1 + (m * m) / (pabs * pabs));
  if (transport) {
    // phi and theta columns
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const V3 seed = mk(T.br[c], T.br[3 + c], T.br[6 + c]);
      const V3 tKick1 = cross(seed, bend1);
      const V3 tDir2 = add3(tKick1, seed);
      const V3 tDir3 = add3(seed, cross(tDir2, bend2));
      const V3 tDir4 = add3(seed, cross(tDir3, bend2));
      const V3 tDirEnd = sub3(scale3(2, tDir4), seed);
      const V3 tKick4 = cross(tDirEnd, bend3);
      const V3 dp = add3(add3(tDir2, tDir3), tDir4);
      const V3 dd = add3(add3(add3(tKick1, scale3(2, tDir3)), tDirEnd), tKick4);
      T.tr[c] = T.tr[c] + hThird * dp.x;
      T.tr[3 + c] = T.tr[3 + c] + hThird * dp.y;
      T.tr[6 + c] = T.tr[6 + c] + hThird * dp.z;
      T.br[c] = dd.x / 3;
      T.br[3 + c] = dd.y / 3;
      T.br[6 + c] = dd.z / 3;
    }
    // q/p column: direction part scaled by q/p plus the stages' own q/p
    // dependence (M(7, 4) = 1 in vacuum)
    {
      const double invQop = 1 / qop;
      const double posW = hThird * invQop, dirW = invQop / 3;
      const V3 seed = scale3(qop, mk(T.br[2], T.br[5], T.br[8]));
      const V3 tKick1 = add3(scale3(1.0, kick1), cross(seed, bend1));
      const V3 tDir2 = add3(tKick1, seed);
      const V3 tDir3 = add3(add3(scale3(1.0, sub3(dir3, dir)), seed), cross(tDir2, bend2));
      const V3 tDir4 = add3(add3(scale3(1.0, sub3(dir4, dir)), seed), cross(tDir3, bend2));
      const V3 tDirEnd = sub3(scale3(2, tDir4), seed);
      const V3 tKick4 = add3(scale3(1.0, kick4), cross(tDirEnd, bend3));
      const V3 dp = add3(add3(tDir2, tDir3), tDir4);
      const V3 dd = add3(add3(add3(tKick1, scale3(2, tDir3)), tDirEnd), tKick4);
      T.tr[2] = T.tr[2] + posW * dp.x;
      T.tr[5] = T.tr[5] + posW * dp.y;
      T.tr[8] = T.tr[8] + posW * dp.z;
      T.br[2] = dirW * dd.x;
      T.br[5] = dirW * dd.y;
      T.br[8] = dirW * dd.z;
      const double dtDqop = h * (m * m) / ((pabs * pabs) * qop * dtds);
      T.tr37 = T.tr37 + dtDqop * 1.0;
    }
    const double twoOverH = 2 / h;
    T.der[0] = newDir.x;
    T.der[1] = newDir.y;
    T.der[2] = newDir.z;
    T.der[3] = dtds;
    T.der[4] = twoOverH * kick4.x;
    T.der[5] = twoOverH * kick4.y;
    T.der[6] = twoOverH * kick4.z;
  }
  T.pos = newPos;
  T.time = T.time + h * dtds;
  T.dir = newDir;
  T.pathAcc += h;
  ++T.nSteps;
  T.nTrials += nTrials;
  const double nextAccuracy = fabs(h * stepScaling(o.stepTolerance, err));
  const double previousAccuracy = fabs(T.cAcc);
  const double initialStepLength = fabs(initialH);
  if (nextAccuracy < initialStepLength || nextAccuracy > previousAccuracy)
    T.cAcc = nextAccuracy;
  return 0;
}

}  // namespace kf
