// kf_oracle.cpp -- CPU restatement of the reference Kalman-fit path.
// TEST INFRASTRUCTURE ONLY (see kf_oracle.hpp).
#include "kf_oracle.hpp"

#include <algorithm>
#include <cmath>

namespace acts_oracle {

// ===========================================================================
// Material.cpp / MaterialSlab.hpp / Interactions.cpp (float arithmetic)
// ===========================================================================

// Units.hpp literals are evaluated as double(unit * long double literal).
static const double kMeV = 1e-3;
static const double kEV = 1e-9;
static const float kMe = static_cast<float>(static_cast<double>(kMeV * 0.5109989461L));
static const float kK = static_cast<float>(static_cast<double>(kMeV * 30.7075L) * 1.0);
static const float kPlasmaEnergyScale = static_cast<float>(static_cast<double>(kEV * 28.816L));
static const double kHighland = static_cast<double>(kMeV * 13.6L);

// Material.cpp: fromMolarDensity, approximateMeanExcitationEnergy
MaterialF materialFromMolarDensity(float x0, float l0, float ar, float z,
                                   float molarRho) {
  MaterialF m;
  m.x0 = x0;
  m.l0 = l0;
  m.ar = ar;
  m.z = z;
  m.molarRho = molarRho;
  m.molarElectronRho = z * molarRho;
  // 16_eV * std::pow(z, 0.9f): double * float, narrowed to the float return
  m.meanI = static_cast<float>(static_cast<double>(kEV * 16.0L) * std::pow(z, 0.9f));
  return m;
}

// MaterialSlab.hpp:125-130
SlabF SlabF::make(const MaterialF& m, float t) {
  SlabF s;
  s.mat = m;
  s.thickness = t;
  const float eps = 2 * std::numeric_limits<float>::epsilon();
  s.tInX0 = (eps < m.x0) ? (t / m.x0) : 0;
  s.tInL0 = (eps < m.l0) ? (t / m.l0) : 0;
  return s;
}

namespace {
// Natural log of a float as the reference's std::log(float): taken as the
// correctly rounded value (glibc's logf differs from it in ~0.02 % of
// arguments, which is libm-version dependent and parity unpinned).
inline float logF(float x) { return static_cast<float>(std::log(static_cast<double>(x))); }
inline float log10F(float x) { return static_cast<float>(std::log10(static_cast<double>(x))); }

// Interactions.cpp:32-58
struct RelQ {
  float q2OverBeta2, beta2, betaGamma, gamma;
  RelQ(float mass, float qOverP, float absQ) {
    q2OverBeta2 = absQ * absQ + (mass * qOverP) * (mass * qOverP);
    const float mOverP = mass * std::abs(qOverP / absQ);
    const float pOverM = 1.0f / mOverP;
    beta2 = 1.0f / (1.0f + mOverP * mOverP);
    betaGamma = pOverM;
    gamma = std::sqrt(1.0f * 1.0f + pOverM * pOverM);
  }
};
inline float massTerm(float mass, const RelQ& rq) {
  return 2 * mass * rq.betaGamma * rq.betaGamma;
}
inline float wMax(float mass, const RelQ& rq) {
  const float mfrac = kMe / mass;
  const float nom = 2 * kMe * rq.betaGamma * rq.betaGamma;
  const float den = 1.0f + 2 * rq.gamma * mfrac + mfrac * mfrac;
  return nom / den;
}
inline float epsilonE(float ne, float thickness, const RelQ& rq) {
  return 0.5f * kK * ne * thickness * rq.q2OverBeta2;
}
inline float deltaHalf(float I, float ne, const RelQ& rq) {
  if (rq.betaGamma < 10.0f) return 0.0f;
  const float plasma = kPlasmaEnergyScale *
                       std::sqrt(ne / static_cast<float>(1 / 1000.0));
  return logF(rq.betaGamma) + logF(plasma / I) - 0.5f;
}
}  // namespace

// Interactions.cpp:174-198
float energyLossBethe(const SlabF& slab, float m, float qOverP, float absQ) {
  if (slab.isVacuum()) return 0.0f;
  const RelQ rq{m, qOverP, absQ};
  const float I = slab.mat.meanI;
  const float ne = slab.mat.molarElectronRho;
  const float eps = epsilonE(ne, slab.thickness, rq);
  const float dhalf = deltaHalf(I, ne, rq);
  const float u = massTerm(kMe, rq);
  const float wmax = wMax(m, rq);
  const float running =
      logF(u / I) + logF(wmax / I) - 2.0f * rq.beta2 - 2.0f * dhalf;
  return eps * running;
}

// Interactions.cpp:237-255
float energyLossLandau(const SlabF& slab, float m, float qOverP, float absQ) {
  if (slab.isVacuum()) return 0.0f;
  const RelQ rq{m, qOverP, absQ};
  const float I = slab.mat.meanI;
  const float ne = slab.mat.molarElectronRho;
  const float eps = epsilonE(ne, slab.thickness, rq);
  const float dhalf = deltaHalf(I, ne, rq);
  const float u = massTerm(kMe, rq);
  const float running =
      logF(u / I) + logF(eps / I) + 0.2f - rq.beta2 - 2 * dhalf;
  return eps * running;
}

// Interactions.cpp:157-168, 307-311
float energyLossLandauFwhm(const SlabF& slab, float m, float qOverP,
                           float absQ) {
  const RelQ rq{m, qOverP, absQ};
  if (slab.isVacuum()) return 0.0f;
  return 4 * epsilonE(slab.mat.molarElectronRho, slab.thickness, rq);
}

// Interactions.cpp:313-329
float energyLossLandauSigmaQOverP(const SlabF& slab, float m, float qOverP,
                                  float absQ) {
  const RelQ rq{m, qOverP, absQ};
  float fwhm = 0.0f;
  if (!slab.isVacuum())
    fwhm = 4 * epsilonE(slab.mat.molarElectronRho, slab.thickness, rq);
  const float sigmaE = fwhm * 0.42466090014400953f;
  const float pInv = qOverP / absQ;
  const float qOverBeta = std::sqrt(rq.q2OverBeta2);
  return qOverBeta * pInv * pInv * sigmaE;
}

// Interactions.cpp:484-530
float multipleScatteringTheta0(const SlabF& slab, int absPdg, float m,
                               float qOverP, float absQ) {
  if (slab.isVacuum()) return 0.0f;
  const float xOverX0 = slab.tInX0;
  const float momentumInv = std::abs(qOverP / absQ);
  const float q2OverBeta2 = RelQ(m, qOverP, absQ).q2OverBeta2;
  const float t = std::sqrt(xOverX0 * q2OverBeta2);
  if (absPdg == 11) {
    const double k175 = static_cast<double>(kMeV * 17.5L);
    return static_cast<float>(k175 * momentumInv * t *
                              (1.0f + 0.125f * log10F(10.0f * xOverX0)));
  }
  return static_cast<float>(kHighland * momentumInv * t *
                            (1.0f + 0.038f * 2 * logF(t)));
}

// ===========================================================================
// periodic.hpp:19-93, TrackParameterHelpers.hpp:51-83
// ===========================================================================
double wrapPeriodic(double value, double start, double range) {
  double diff = value - start;
  return ((0 <= diff) && (diff < range))
             ? value
             : (value - range * std::floor(diff / range));
}
double radianPos(double x) { return wrapPeriodic(x, 0.0, 2 * kPi); }
double radianSym(double x) { return wrapPeriodic(x, -kPi, 2 * kPi); }
void normalizePhiTheta(double& phi, double& theta) {
  theta = radianPos(theta);
  if (kPi < theta) {
    phi += kPi;
    theta = 2 * kPi - theta;
  }
  phi = radianSym(phi);
}

// ===========================================================================
// Surfaces
// ===========================================================================

// JacobianHelpers.hpp:26-64 with VectorHelpers.hpp:165-180
static void trig(const V3& d, double& cosPhi, double& sinPhi, double& cosTheta,
                 double& sinTheta) {
  cosTheta = d[2];
  sinTheta = fastCathetus1(1, d[2]);
  const double inv = 1. / sinTheta;
  cosPhi = d[0] * inv;
  sinPhi = d[1] * inv;
}

// RealQuadraticEquation.hpp:57-66
struct Quadratic {
  double first = 0, second = 0;
  int solutions = 0;
  Quadratic(double alpha, double beta, double gamma) {
    double disc = beta * beta - 4 * alpha * gamma;
    if (disc >= 0) {
      solutions = (disc == 0) ? 1 : 2;
      double q = -0.5 * (beta + (beta > 0 ? std::sqrt(disc) : -std::sqrt(disc)));
      first = q / alpha;
      second = gamma / q;
    }
  }
};

MultiIntersection Surface::intersect(const V3& pos, const V3& dir,
                                     double tol) const {
  MultiIntersection mi;
  if (type == ACTS_KF_SURFACE_LINE) {
    // LineSurface.cpp:136-186: closest approach of the straight line
    // (pos, dir) to the surface's axis; no bounds (PerigeeSurface)
    mi.size = 1;
    const V3 mb = v3(pl.t[0], pl.t[1], pl.t[2]);
    const V3 eb = pl.col(2);
    const V3 mab = v3(mb[0] - pos[0], mb[1] - pos[1], mb[2] - pos[2]);
    const double eaTeb = dot(dir, eb);
    const double denom = 1 - eaTeb * eaTeb;
    if (std::abs(denom) < std::abs(tol)) return mi;  // invalid
    const double u = (dot(mab, dir) - dot(mab, eb) * eaTeb) / denom;
    mi.is[0].path = u;
    mi.is[0].status = std::abs(u) > std::abs(tol) ? IStatus::reachable : IStatus::onSurface;
    return mi;
  }
  if (type == ACTS_KF_SURFACE_CYLINDER) {
    // CylinderSurface.cpp:198-283 (boundary tolerance infinite)
    const double Rr = b0;
    V3 caxis = pl.col(2);
    V3 ccenter = v3(pl.t[0], pl.t[1], pl.t[2]);
    V3 pc = v3(pos[0] - ccenter[0], pos[1] - ccenter[1], pos[2] - ccenter[2]);
    V3 pcXcd = cross(pc, caxis);
    V3 ldXcd = cross(dir, caxis);
    double a = dot(ldXcd, ldXcd);
    double b = 2. * dot(ldXcd, pcXcd);
    double c = dot(pcXcd, pcXcd) - (Rr * Rr);
    Quadratic qe(a, b, c);
    mi.size = 2;
    if (qe.solutions == 0) return mi;  // two invalid
    Intersection first;
    first.path = qe.first;
    first.status = std::abs(qe.first) < std::abs(tol) ? IStatus::onSurface
                                                      : IStatus::reachable;
    if (qe.solutions == 1) {
      mi.is[0] = first;
      mi.is[1] = first;
      return mi;
    }
    Intersection second;
    second.path = qe.second;
    second.status = std::abs(qe.second) < std::abs(tol) ? IStatus::onSurface
                                                        : IStatus::reachable;
    if (first.path <= second.path) {
      mi.is[0] = first;
      mi.is[1] = second;
    } else {
      mi.is[0] = second;
      mi.is[1] = first;
    }
    return mi;
  }
  // PlanarHelper.hpp:56-77 (plane and disc, boundary tolerance infinite)
  mi.size = 1;
  V3 n = pl.col(2);
  double denom = dot(dir, n);
  if (denom == 0) return mi;
  V3 d = v3(pl.t[0] - pos[0], pl.t[1] - pos[1], pl.t[2] - pos[2]);
  double path = dot(n, d) / denom;
  mi.is[0].path = path;
  mi.is[0].status =
      std::abs(path) < std::abs(tol) ? IStatus::onSurface : IStatus::reachable;
  return mi;
}

bool Surface::globalToLocal(const V3& pos, double& l0, double& l1,
                            double tol) const {
  V3 loc = pl.toLocal(pos);
  if (type == ACTS_KF_SURFACE_CYLINDER) {
    // CylinderSurface.cpp:125-145
    double inttol = tol;
    if (tol == kOnSurfaceTolerance) inttol = b0 * 0.0001;
    if (inttol < 0.01) inttol = 0.01;
    if (std::abs(perp(loc) - b0) > inttol) return false;
    l0 = b0 * phiOf(loc);
    l1 = loc[2];
    return true;
  }
  if (std::abs(loc[2]) > std::abs(tol)) return false;
  if (type == ACTS_KF_SURFACE_DISC) {
    // DiscSurface.cpp:91-100
    l0 = perp(loc);
    l1 = phiOf(loc);
  } else {
    // PlaneSurface.cpp:77-85
    l0 = loc[0];
    l1 = loc[1];
  }
  return true;
}

V3 Surface::localToGlobal(double l0, double l1) const {
  if (type == ACTS_KF_SURFACE_CYLINDER) {
    // CylinderSurface.cpp:116-123
    double r = b0;
    double phi = l0 / r;
    return pl.toGlobal(v3(r * std::cos(phi), r * std::sin(phi), l1));
  }
  if (type == ACTS_KF_SURFACE_DISC) {
    // DiscSurface.cpp:82-89
    return pl.toGlobal(v3(l0 * std::cos(l1), l0 * std::sin(l1), 0.));
  }
  // PlaneSurface.cpp:72-75
  return pl.toGlobal(v3(l0, l1, 0.));
}

// CylinderSurface.cpp:158-167
V3 Surface::cylinderNormal(const V3& pos) const {
  V3 p3 = pl.toLocal(pos);
  p3[2] = 0.;
  return pl.rotate(normalized(p3));
}

Mat<3, 3> Surface::referenceFrame(const V3& pos) const {
  Mat<3, 3> f;
  if (type == ACTS_KF_SURFACE_CYLINDER) {
    // CylinderSurface.cpp:93-110
    V3 measY = pl.col(2);
    V3 measDepth = cylinderNormal(pos);
    V3 measX = normalized(cross(measY, measDepth));
    for (int i = 0; i < 3; ++i) {
      f.a[i][0] = measX[i];
      f.a[i][1] = measY[i];
      f.a[i][2] = measDepth[i];
    }
    return f;
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) f.a[i][j] = pl.R[i][j];
  return f;
}

// LineSurface.cpp:101-114; other types: referenceFrame(pos)
Mat<3, 3> Surface::referenceFrame(const V3& pos, const V3& dir) const {
  if (type != ACTS_KF_SURFACE_LINE) return referenceFrame(pos);
  const V3 unitZ0 = pl.col(2);
  const V3 unitD0 = normalized(cross(unitZ0, dir));
  const V3 unitDistance = cross(unitD0, unitZ0);
  Mat<3, 3> f;
  for (int i = 0; i < 3; ++i) {
    f.a[i][0] = unitD0[i];
    f.a[i][1] = unitZ0[i];
    f.a[i][2] = unitDistance[i];
  }
  return f;
}

// LineSurface.cpp:68-96: the frame's inverse applied to pos - center.  The
// frame is orthonormal, so its inverse is its transpose (Eigen inverts the
// 3x3 by cofactors; the two agree up to the last bits).
bool Surface::globalToLocal(const V3& pos, const V3& dir, double& l0, double& l1,
                            double tol) const {
  if (type != ACTS_KF_SURFACE_LINE) return globalToLocal(pos, l0, l1, tol);
  const Mat<3, 3> f = referenceFrame(pos, dir);
  const V3 d = v3(pos[0] - pl.t[0], pos[1] - pl.t[1], pos[2] - pl.t[2]);
  double loc[3];
  for (int j = 0; j < 3; ++j) loc[j] = (f.a[0][j] * d[0] + f.a[1][j] * d[1]) + f.a[2][j] * d[2];
  if (std::abs(loc[2]) > std::abs(tol)) return false;
  l0 = loc[0];
  l1 = loc[1];
  return true;
}

// LineSurface.cpp:57-66
V3 Surface::localToGlobal(double l0, double l1, const V3& dir) const {
  if (type != ACTS_KF_SURFACE_LINE) return localToGlobal(l0, l1);
  const V3 unitZ0 = pl.col(2);
  const V3 radiusAxis = normalized(cross(unitZ0, dir));
  const V3 locZ = pl.toGlobal(v3(0., 0., l1));
  return v3(locZ[0] + l0 * radiusAxis[0], locZ[1] + l0 * radiusAxis[1],
            locZ[2] + l0 * radiusAxis[2]);
}

double Surface::pathCorrection(const V3& pos, const V3& dir) const {
  if (type == ACTS_KF_SURFACE_CYLINDER) {
    // CylinderSurface.cpp:169-175
    V3 n = cylinderNormal(pos);
    double cosAlpha = dot(n, dir);
    return std::abs(1. / cosAlpha);
  }
  // PlaneSurface.cpp:161-166, DiscSurface.cpp:367-372
  return 1. / std::abs(dot(pl.col(2), dir));
}

Mat<8, 6> Surface::boundToFreeJacobian(const V3& pos, const V3& dir) const {
  Mat<8, 6> J = Mat<8, 6>::zero();
  double cosPhi, sinPhi, cosTheta, sinTheta;
  trig(dir, cosPhi, sinPhi, cosTheta, sinTheta);
  if (type == ACTS_KF_SURFACE_DISC) {
    // DiscSurface.cpp:197-228
    V3 posLoc = pl.toLocal(pos);
    const double lr = perp(posLoc);
    const double lphi = phiOf(posLoc);
    const double lc = std::cos(lphi), ls = std::sin(lphi);
    for (int i = 0; i < 3; ++i) {
      const double lx = pl.R[i][0], ly = pl.R[i][1];
      J.a[i][0] = lc * lx + ls * ly;
      J.a[i][1] = lr * (lc * ly - ls * lx);
    }
  } else {
    // Surface.cpp:249-268
    Mat<3, 3> f = referenceFrame(pos, dir);
    for (int i = 0; i < 3; ++i) {
      J.a[i][0] = f.a[i][0];
      J.a[i][1] = f.a[i][1];
    }
  }
  J.a[3][5] = 1;
  J.a[4][2] = -dir[1];
  J.a[4][3] = cosTheta * cosPhi;
  J.a[5][2] = dir[0];
  J.a[5][3] = cosTheta * sinPhi;
  J.a[6][2] = 0;
  J.a[6][3] = -sinTheta;
  J.a[7][4] = 1;
  if (type == ACTS_KF_SURFACE_LINE) {
    // LineSurface.cpp:188-225: the position moves with the direction
    const Mat<3, 3> f = referenceFrame(pos, dir);
    double l0 = 0, l1 = 0;
    globalToLocal(pos, dir, l0, l1, std::numeric_limits<double>::max());
    const V3 fx = v3(f.a[0][0], f.a[1][0], f.a[2][0]);
    const V3 fy = v3(f.a[0][1], f.a[1][1], f.a[2][1]);
    const V3 fz = v3(f.a[0][2], f.a[1][2], f.a[2][2]);
    const double ipdn = 1. / dot(dir, fz);
    V3 dDPhiY = cross(fy, v3(J.a[4][2], J.a[5][2], J.a[6][2]));
    V3 dDThetaY = cross(fy, v3(J.a[4][3], J.a[5][3], J.a[6][3]));
    const double sp = dot(fx, dDPhiY), st = dot(fx, dDThetaY);
    for (int i = 0; i < 3; ++i) {
      dDPhiY[i] = dDPhiY[i] - fx[i] * sp;
      dDThetaY[i] = dDThetaY[i] - fx[i] * st;
    }
    for (int i = 0; i < 3; ++i) {
      J.a[i][2] = dDPhiY[i] * l0 * ipdn;
      J.a[i][3] = dDThetaY[i] * l0 * ipdn;
    }
  }
  return J;
}

Mat<6, 8> Surface::freeToBoundJacobian(const V3& pos, const V3& dir) const {
  Mat<6, 8> J = Mat<6, 8>::zero();
  double cosPhi, sinPhi, cosTheta, sinTheta;
  trig(dir, cosPhi, sinPhi, cosTheta, sinTheta);
  const double invSinTheta = 1. / sinTheta;
  if (type == ACTS_KF_SURFACE_DISC) {
    // DiscSurface.cpp:230-264
    V3 posLoc = pl.toLocal(pos);
    const double lr = perp(posLoc);
    const double lphi = phiOf(posLoc);
    const double lc = std::cos(lphi), ls = std::sin(lphi);
    for (int k = 0; k < 3; ++k) {
      const double lx = pl.R[k][0], ly = pl.R[k][1];
      J.a[0][k] = lc * lx + ls * ly;
      J.a[1][k] = (lc * ly - ls * lx) / lr;
    }
  } else {
    // Surface.cpp:270-290
    Mat<3, 3> f = referenceFrame(pos, dir);
    for (int k = 0; k < 3; ++k) {
      J.a[0][k] = f.a[k][0];
      J.a[1][k] = f.a[k][1];
    }
  }
  J.a[5][3] = 1;
  J.a[2][4] = -sinPhi * invSinTheta;
  J.a[2][5] = cosPhi * invSinTheta;
  J.a[2][6] = 0;
  J.a[3][4] = cosPhi * cosTheta;
  J.a[3][5] = sinPhi * cosTheta;
  J.a[3][6] = -sinTheta;
  J.a[4][7] = 1;
  return J;
}

// Surface.cpp:292-307; LineSurface.cpp:227-250
Mat<1, 8> Surface::freeToPathDerivative(const V3& pos, const V3& dir) const {
  if (type == ACTS_KF_SURFACE_LINE) {
    const V3 pc = v3(pos[0] - pl.t[0], pos[1] - pl.t[1], pos[2] - pl.t[2]);
    const V3 z = pl.col(2);
    const double pz = dot(pc, z);
    const double dz = dot(z, dir);
    const double norm = 1 / (1 - dz * dz);
    Mat<1, 8> s = Mat<1, 8>::zero();
    for (int i = 0; i < 3; ++i) {
      s.a[0][i] = norm * (dz * z[i] - dir[i]);
      s.a[0][4 + i] = norm * (pz * z[i] - pc[i]);
    }
    return s;
  }
  Mat<3, 3> f = referenceFrame(pos);
  V3 z = v3(f.a[0][2], f.a[1][2], f.a[2][2]);
  const double dz = dot(z, dir);
  Mat<1, 8> s = Mat<1, 8>::zero();
  for (int i = 0; i < 3; ++i) s.a[0][i] = (-1.0 * z[i]) / dz;
  return s;
}

// TransformationHelpers.cpp:17-40
Vec8 boundToFree(const Surface& s, const Vec6& b) {
  const double sinTheta = std::sin(b.a[3][0]);
  V3 dir = v3(std::cos(b.a[2][0]) * sinTheta, std::sin(b.a[2][0]) * sinTheta,
              std::cos(b.a[3][0]));
  V3 pos = s.localToGlobal(b.a[0][0], b.a[1][0], dir);
  Vec8 f;
  f.a[0][0] = pos[0];
  f.a[1][0] = pos[1];
  f.a[2][0] = pos[2];
  f.a[3][0] = b.a[5][0];
  f.a[4][0] = dir[0];
  f.a[5][0] = dir[1];
  f.a[6][0] = dir[2];
  f.a[7][0] = b.a[4][0];
  return f;
}

// TransformationHelpers.cpp:42-63
int freeToBound(const Vec8& f, const Surface& s, Vec6& out) {
  V3 pos = v3(f.a[0][0], f.a[1][0], f.a[2][0]);
  V3 dir = v3(f.a[4][0], f.a[5][0], f.a[6][0]);
  double l0, l1;
  if (!s.globalToLocal(pos, dir, l0, l1)) return ACTS_KF_STATUS_NOT_ON_SURFACE;
  out.a[0][0] = l0;
  out.a[1][0] = l1;
  out.a[5][0] = f.a[3][0];
  out.a[2][0] = phiOf(dir);
  out.a[3][0] = thetaOf(dir);
  out.a[4][0] = f.a[7][0];
  return 0;
}

// ===========================================================================
// Field
// ===========================================================================
// fieldMapXYZ (BFieldMapUtils.cpp:109-186): axis a has nb[a] bins of width
// (max - min) / nb on [min, max) with max one step past the last point
// (Helpers.hpp:289-310, Axis.hpp:45); bin b (1-based) has lower edge
// min + (b - 1) * width and holds the value of point b - 1.
void FieldGrid::set(const int n[3], const double pmin[3], const double pmax[3],
                    const double* values) {
  for (int a = 0; a < 3; ++a) {
    nb[a] = n[a];
    min[a] = pmin[a];
    const double step = (pmax[a] - pmin[a]) / static_cast<double>(n[a] - 1);
    max[a] = pmax[a] + step;
    width[a] = (max[a] - min[a]) / static_cast<double>(n[a]);
  }
  v.assign(values, values + 3 * static_cast<size_t>(n[0]) * n[1] * n[2]);
}

// InterpolatedBFieldMap::getField (InterpolatedBFieldMap.hpp:335-366):
// inside iff lowerLeft <= x < lower edge of the last bin (:223-229, 298-310);
// the cell of the point's bin; Acts::interpolate (Interpolation.hpp:96-120)
// collapses the last axis first with (1 - f) * a + f * b.
int FieldGrid::get(const V3& pos, V3& out) const {
  int lo[3];
  double ll[3], ur[3];
  for (int a = 0; a < 3; ++a) {
    const double upper = min[a] + static_cast<double>(nb[a] - 1) * width[a];
    // (a NaN coordinate, undefined in the reference, counts as outside)
    if (!(pos[a] >= min[a]) || pos[a] >= upper) return ACTS_KF_STATUS_FIELD_OUT_OF_BOUNDS;
    // Axis::getBin (Axis.hpp:232-235), 1-based, here inside [1, nb - 1]
    const int bin = static_cast<int>(std::floor((pos[a] - min[a]) / width[a]) + 1);
    lo[a] = bin - 1;  // grid point of the lower corner
    ll[a] = min[a] + static_cast<double>(bin - 1) * width[a];
    ur[a] = min[a] + static_cast<double>(bin) * width[a];
  }
  // corners in the canonical order: bit 2 = x, bit 1 = y, bit 0 = z
  double c[8][3];
  for (int i = 0; i < 8; ++i) {
    const int ix = lo[0] + ((i >> 2) & 1), iy = lo[1] + ((i >> 1) & 1), iz = lo[2] + (i & 1);
    // rounding can put a point just below the last grid point into the last
    // bin, whose upper corner is the overflow bin: zero (setExteriorBins)
    if (ix >= nb[0] || iy >= nb[1] || iz >= nb[2]) {
      c[i][0] = c[i][1] = c[i][2] = 0.0;
      continue;
    }
    const double* p = &v[3 * ((static_cast<size_t>(ix) * nb[1] + iy) * nb[2] + iz)];
    c[i][0] = p[0];
    c[i][1] = p[1];
    c[i][2] = p[2];
  }
  int n = 8;
  for (int a = 2; a >= 0; --a) {
    const double f = (pos[a] - ll[a]) / (ur[a] - ll[a]);
    for (int i = 0; i < n / 2; ++i)
      for (int k = 0; k < 3; ++k) c[i][k] = (1 - f) * c[2 * i][k] + f * c[2 * i + 1][k];
    n /= 2;
  }
  out = v3(c[0][0], c[0][1], c[0][2]);
  return 0;
}

int Field::get(const V3& pos, V3& out) const {
  if (kind == ACTS_KF_FIELD_CONSTANT) {
    out = v3(B[0], B[1], B[2]);
    return 0;
  }
  if (kind == ACTS_KF_FIELD_GRID_XYZ) return grid.get(pos, out);
  return ACTS_KF_STATUS_FIELD_OUT_OF_BOUNDS;
}

// ===========================================================================
// SolenoidBField (SolenoidBField.cpp:24-162), used to sample field maps.
// The reference passes k^2 to Boost's ellint_1/ellint_2, which take the
// modulus: the integrals evaluated are K and E of modulus k^2.
// ===========================================================================
// complete elliptic integrals K(k), E(k) of modulus k by the arithmetic-
// geometric mean (Abramowitz & Stegun 17.6)
void ellipticKE(double k, double& K, double& E) {
  double a = 1.0, b = std::sqrt(1.0 - k * k), c = k;
  double sum = 0.5 * c * c, pow2 = 0.5;
  // stop once c has reached the rounding level of a: further terms are
  // below 1e-30, and a - b only oscillates in the last bit from there on
  for (int i = 0; i < 64 && std::fabs(c) > 1e-15 * a; ++i) {
    const double an = 0.5 * (a + b);
    const double bn = std::sqrt(a * b);
    c = 0.5 * (a - b);
    a = an;
    b = bn;
    pow2 *= 2.0;
    sum += pow2 * c * c;
  }
  K = kPi / (2.0 * a);
  E = K * (1.0 - sum);
}

Solenoid::Solenoid(double radius_, double length_, int nCoils_, double bMagCenter)
    : radius(radius_), length(length_), nCoils(nCoils_) {
  dz = length / nCoils;
  R2 = radius * radius;
  scale = 1.0;
  double br, bz;
  multiCoil(0.0, 0.0, 1.0, br, bz);
  scale = bMagCenter / std::sqrt(br * br + bz * bz);
}

double Solenoid::k2(double r, double z) const {
  return 4 * radius * r / ((radius + r) * (radius + r) + z * z);
}

double Solenoid::Br(double rs, double z, double sc) const {
  const double r = std::abs(rs);
  if (r == 0) return 0.;
  const double k_2 = k2(r, z);
  const double k = std::sqrt(k_2);
  const double constant = sc * k * z / (4 * kPi * std::sqrt(radius * r * r * r));
  double K, E;
  ellipticKE(k_2, K, E);
  const double B = (2. - k_2) / (2. - 2. * k_2) * E - K;
  return r / rs * constant * B;
}

double Solenoid::Bz(double rs, double z, double sc) const {
  const double r = std::abs(rs);
  if (r == 0) return sc / 2. * R2 / (std::sqrt(R2 + z * z) * (R2 + z * z));
  const double k_2 = k2(r, z);
  const double k = std::sqrt(k_2);
  const double constant = sc * k / (4 * kPi * std::sqrt(radius * r));
  double K, E;
  ellipticKE(k_2, K, E);
  const double B = ((radius + r) * k_2 - 2. * r) / (2. * r * (1. - k_2)) * E + K;
  return constant * B;
}

void Solenoid::multiCoil(double r, double z, double sc, double& br, double& bz) const {
  br = 0;
  bz = 0;
  for (int coil = 0; coil < nCoils; ++coil) {
    const double zs = z + length * 0.5 - dz * (coil + 0.5);
    br += Br(r, zs, sc);
    bz += Bz(r, zs, sc);
  }
}

V3 Solenoid::field(const V3& pos) const {
  const double r = perp(pos);
  double br, bz;
  multiCoil(r, pos[2], scale, br, bz);
  V3 out = v3(0, 0, bz);
  if (r != 0.) {
    const V3 rDir = normalized(v3(pos[0], pos[1], 0));
    out = v3(out[0] + rDir[0] * br, out[1] + rDir[1] * br, out[2] + rDir[2] * br);
  }
  return out;
}

// ===========================================================================
// EigenStepper state + ConstrainedStep
// ===========================================================================
namespace {

// ConstrainedStep.hpp:44-173
struct ConstrainedStep {
  enum Type { Navigator = 0, Actor = 1, User = 2 };
  double values[3] = {kNotSet, kNotSet, kNotSet};
  double accuracy = kNotSet;
  double value() const {
    double mn = std::min(std::min(values[0], values[1]), values[2]);
    double r = std::min(std::abs(mn), accuracy);
    return std::signbit(mn) ? -r : r;
  }
  double value(Type t) const { return values[t]; }
  void release(Type t) { values[t] = kNotSet; }
  void update(double v, Type t) {
    if (std::abs(v) < std::abs(values[t])) values[t] = v;
  }
};

// EigenStepper.hpp:81-153 (covariance transport always on for the fit)
struct Stepper {
  Vec8 pars;
  bool covTransport = false;
  M6 cov = M6::zero();
  M6 jacobian = M6::identity();
  Mat<8, 6> jacToGlobal = Mat<8, 6>::zero();
  M8 jacTransport = M8::identity();
  Vec8 derivative = Vec8::zero();
  double pathAccumulated = 0;
  int nSteps = 0, nStepTrials = 0;
  ConstrainedStep stepSize;
  double previousStepSize = 0;
  V3 Bfirst, Bmiddle, Blast, k1, k2, k3, k4;

  V3 position() const { return v3(pars.a[0][0], pars.a[1][0], pars.a[2][0]); }
  V3 direction() const { return v3(pars.a[4][0], pars.a[5][0], pars.a[6][0]); }
  double qop() const { return pars.a[7][0]; }
  void updateStepSize(double s, ConstrainedStep::Type t) {
    previousStepSize = stepSize.value();
    stepSize.update(s, t);
  }
};

struct StepResult {
  int status;
  double h;
};

// EigenStepperDefaultExtension.hpp:40-58
inline V3 kFirst(double qop, const V3& dir, const V3& B) {
  V3 c = cross(dir, B);
  return v3(qop * c[0], qop * c[1], qop * c[2]);
}
inline V3 kNext(double qop, const V3& dir, double h, const V3& kprev,
                const V3& B) {
  V3 d = v3(dir[0] + h * kprev[0], dir[1] + h * kprev[1], dir[2] + h * kprev[2]);
  V3 c = cross(d, B);
  return v3(qop * c[0], qop * c[1], qop * c[2]);
}

inline Mat<3, 3> crossCols(const Mat<3, 3>& m, const V3& v) {
  Mat<3, 3> r;
  for (int j = 0; j < 3; ++j) {
    V3 c = cross(v3(m.a[0][j], m.a[1][j], m.a[2][j]), v);
    for (int i = 0; i < 3; ++i) r.a[i][j] = c[i];
  }
  return r;
}

// SympyStepper::step (SympyStepper.cpp:141-305) with the vacuum kernel
// rk4_vacuum that generate_sympy_stepper.py writes (its
// _atlas_rk4_stages and rk4_vacuum_b2f_atlasexpr, :157-444): an ATLAS-form
// RK4 step over half-step "bend" vectors b_i = (h q/p / 2) B_i, whose
// error estimate is 2|h| |(2d + h k1/2 + h k4/2) - (dir3 + dir4)|_1 = h^2
// |k1 - k2 - k3 + k4|_1, and which carries the bound-to-free jacobian M
// (here st.jacToGlobal; st.jacTransport stays the identity) through the
// same recursion: the phi, theta and q/p columns of M as tangents of the
// stages.  The generated code's operation order is sympy's and is not
// available here; this is the recursion in its natural order.
StepResult stepSympy(Stepper& st, const Field& field, const Particle& ph,
                     const acts_kf_options& o, double propDir) {
  const V3 pos = st.position();
  const V3 dir = st.direction();
  const double t = st.pars.a[3][0];
  const double qop = st.qop();
  const double pabs = ph.extractMomentum(qop);
  const double m = static_cast<float>(ph.mass);
  auto add = [](const V3& a, const V3& b) { return v3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); };
  auto sub = [](const V3& a, const V3& b) { return v3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); };
  auto scale = [](double s, const V3& a) { return v3(s * a[0], s * a[1], s * a[2]); };
  auto scaling = [&](double err) {
    double x = o.step_tolerance / err;
    x = std::sqrt(std::sqrt(x));
    return std::clamp(x, 0.25, 4.0);
  };
  V3 B1;
  if (int fs = field.get(pos, B1)) return {fs, 0};

  const double initialH = st.stepSize.value() * propDir;
  double h = initialH;
  double err = 0;
  unsigned nTrials = 0;
  V3 bend1, bend2, bend3, kick1, dir2, dir3, dir4, dirEnd, kick4, dirHalfSum;
  double hThird = 0;
  while (true) {
    ++nTrials;
    const double halfHQop = h * qop / 2;
    hThird = h / 3;
    const double hQuarter = h / 4;
    bend1 = scale(halfHQop, B1);
    kick1 = cross(dir, bend1);
    dir2 = add(kick1, dir);
    dirHalfSum = add(dir2, dir);
    const V3 pos2 = add(pos, scale(hQuarter, dirHalfSum));
    V3 B2;
    if (int fs = field.get(pos2, B2)) return {fs, 0};
    bend2 = scale(halfHQop, B2);
    dir3 = add(dir, cross(dir2, bend2));
    dir4 = add(dir, cross(dir3, bend2));
    dirEnd = sub(scale(2, dir4), dir);
    const V3 pos3 = add(pos, scale(h, dir4));
    V3 B3;
    if (int fs = field.get(pos3, B3)) return {fs, 0};
    bend3 = scale(halfHQop, B3);
    kick4 = cross(dirEnd, bend3);
    const V3 e = sub(add(dirHalfSum, kick4), add(dir3, dir4));
    err = 2 * std::abs(h) * ((std::abs(e[0]) + std::abs(e[1])) + std::abs(e[2]));
    const bool ok = !(err > 4.0 * o.step_tolerance);
    err = std::max(1e-20, err);
    if (ok) break;
    h *= scaling(err);
    if (std::abs(h) < std::abs(o.step_size_cutoff))
      return {ACTS_KF_STATUS_STEP_SIZE_STALLED, 0};
    if (nTrials > o.max_runge_kutta_step_trials)
      return {ACTS_KF_STATUS_STEP_ADJUSTMENT_FAILED, 0};
  }
  const V3 newPos = add(pos, scale(hThird, add(add(dir2, dir3), dir4)));
  const V3 newDirX3 = add(scale(2, dir3), add(add(kick1, dirEnd), kick4));
  const double invNorm = 1 / std::sqrt(squaredNorm(newDirX3));
  const V3 newDir = scale(invNorm, newDirX3);
  const double dtds = std::sqrt(1 + (m * m) / (pabs * pabs));
  const double newTime = t + h * dtds;
  if (st.covTransport) {
    Mat<8, 6>& M = st.jacToGlobal;
    // the phi and theta columns: tangents of the stages along the column's
    // direction part
    for (int c = 2; c <= 3; ++c) {
      const V3 seed = v3(M.a[4][c], M.a[5][c], M.a[6][c]);
      const V3 tKick1 = cross(seed, bend1);
      const V3 tDir2 = add(tKick1, seed);
      const V3 tDir3 = add(seed, cross(tDir2, bend2));
      const V3 tDir4 = add(seed, cross(tDir3, bend2));
      const V3 tDirEnd = sub(scale(2, tDir4), seed);
      const V3 tKick4 = cross(tDirEnd, bend3);
      const V3 dp = add(add(tDir2, tDir3), tDir4);
      const V3 dd = add(add(add(tKick1, scale(2, tDir3)), tDirEnd), tKick4);
      for (int i = 0; i < 3; ++i) {
        M.a[i][c] = M.a[i][c] + hThird * dp[i];
        M.a[4 + i][c] = dd[i] / 3;
      }
    }
    // the q/p column: direction part scaled by q/p, plus the stages' own
    // q/p dependence (the bend vectors are linear in q/p)
    {
      const double m74 = M.a[7][4];
      const double invQop = 1 / qop;
      const double posW = hThird * invQop, dirW = invQop / 3;
      const V3 seed = scale(qop, v3(M.a[4][4], M.a[5][4], M.a[6][4]));
      const V3 tKick1 = add(scale(m74, kick1), cross(seed, bend1));
      const V3 tDir2 = add(tKick1, seed);
      const V3 tDir3 = add(add(scale(m74, sub(dir3, dir)), seed), cross(tDir2, bend2));
      const V3 tDir4 = add(add(scale(m74, sub(dir4, dir)), seed), cross(tDir3, bend2));
      const V3 tDirEnd = sub(scale(2, tDir4), seed);
      const V3 tKick4 = add(scale(m74, kick4), cross(tDirEnd, bend3));
      const V3 dp = add(add(tDir2, tDir3), tDir4);
      const V3 dd = add(add(add(tKick1, scale(2, tDir3)), tDirEnd), tKick4);
      for (int i = 0; i < 3; ++i) {
        M.a[i][4] = M.a[i][4] + posW * dp[i];
        M.a[4 + i][4] = dirW * dd[i];
      }
      // time row: d(time)/d(q/p) = h m^2 / (p^2 q/p dt/ds)
      const double dtDqop = h * (m * m) / ((pabs * pabs) * qop * dtds);
      M.a[3][4] = M.a[3][4] + dtDqop * m74;
    }
    // path derivatives: direction, dt/ds, k4 = (2/h) kick4, 0
    const double twoOverH = 2 / h;
    for (int i = 0; i < 3; ++i) st.derivative.a[i][0] = newDir[i];
    st.derivative.a[3][0] = dtds;
    for (int i = 0; i < 3; ++i) st.derivative.a[4 + i][0] = twoOverH * kick4[i];
    st.derivative.a[7][0] = 0;
  }
  for (int i = 0; i < 3; ++i) st.pars.a[i][0] = newPos[i];
  st.pars.a[3][0] = newTime;
  for (int i = 0; i < 3; ++i) st.pars.a[4 + i][0] = newDir[i];
  st.pathAccumulated += h;
  ++st.nSteps;
  st.nStepTrials += nTrials;
  const double sc = scaling(err);
  const double nextAccuracy = std::abs(h * sc);
  const double previousAccuracy = std::abs(st.stepSize.accuracy);
  const double initialStepLength = std::abs(initialH);
  if (nextAccuracy < initialStepLength || nextAccuracy > previousAccuracy)
    st.stepSize.accuracy = nextAccuracy;
  return {0, h};
}

// EigenStepper.ipp:169-388 with EigenStepperDefaultExtension.hpp:104-228
StepResult step(Stepper& st, const Field& field, const Particle& ph,
                const acts_kf_options& o, double propDir) {
  if (o.stepper == ACTS_KF_STEPPER_SYMPY) return stepSympy(st, field, ph, o, propDir);
  double errorEstimate = 0, h2 = 0, half_h = 0;
  const V3 pos = st.position();
  const V3 dir = st.direction();
  const double qop = st.qop();

  int fs = field.get(pos, st.Bfirst);
  if (fs) return {fs, 0};
  st.k1 = kFirst(qop, dir, st.Bfirst);

  auto scaling = [&](double err) {
    double x = o.step_tolerance / err;
    x = std::sqrt(std::sqrt(x));
    return std::clamp(x, 0.25, 4.0);
  };

  auto tryStep = [&](double h, bool& ok) -> int {
    h2 = h * h;
    half_h = h * 0.5;
    V3 pos1;
    for (int i = 0; i < 3; ++i)
      pos1[i] = (pos[i] + half_h * dir[i]) + (h2 * 0.125) * st.k1[i];
    int s = field.get(pos1, st.Bmiddle);
    if (s) return s;
    st.k2 = kNext(qop, dir, half_h, st.k1, st.Bmiddle);
    st.k3 = kNext(qop, dir, half_h, st.k2, st.Bmiddle);
    V3 pos2;
    for (int i = 0; i < 3; ++i)
      pos2[i] = (pos[i] + h * dir[i]) + (h2 * 0.5) * st.k3[i];
    s = field.get(pos2, st.Blast);
    if (s) return s;
    st.k4 = kNext(qop, dir, h, st.k3, st.Blast);
    double l1 = 0;
    for (int i = 0; i < 3; ++i) {
      double v = ((st.k1[i] - st.k2[i]) - st.k3[i]) + st.k4[i];
      l1 = (i == 0) ? std::abs(v) : l1 + std::abs(v);
    }
    errorEstimate = h2 * (l1 + std::abs(0.0 - 0.0 - 0.0 + 0.0));
    errorEstimate = std::max(1e-20, errorEstimate);
    ok = errorEstimate <= 4.0 * o.step_tolerance;
    return 0;
  };

  const double initialH = st.stepSize.value() * propDir;
  double h = initialH;
  unsigned nTrials = 0;
  while (true) {
    ++nTrials;
    bool ok = false;
    int s = tryStep(h, ok);
    if (s) return {s, 0};
    if (ok) break;
    h *= scaling(errorEstimate);
    if (std::abs(h) < std::abs(o.step_size_cutoff))
      return {ACTS_KF_STATUS_STEP_SIZE_STALLED, 0};
    if (nTrials > o.max_runge_kutta_step_trials)
      return {ACTS_KF_STATUS_STEP_ADJUSTMENT_FAILED, 0};
  }

  // propagateTime (EigenStepperDefaultExtension.hpp:112-125)
  const float m = ph.mass;
  const double p = ph.extractMomentum(qop);
  const double dtds = std::sqrt(1 + (m / p) * (m / p));
  st.pars.a[3][0] += h * dtds;
  if (st.covTransport) {
    st.derivative.a[3][0] = dtds;
    // transportMatrix (EigenStepperDefaultExtension.hpp:137-228)
    M8 D = M8::identity();
    V3 dk1dL = cross(dir, st.Bfirst);
    V3 t, c1, c2;
    V3 dk2dL, dk3dL, dk4dL;
    for (int i = 0; i < 3; ++i) t[i] = dir[i] + half_h * st.k1[i];
    c1 = cross(t, st.Bmiddle);
    c2 = cross(dk1dL, st.Bmiddle);
    for (int i = 0; i < 3; ++i) dk2dL[i] = c1[i] + (qop * half_h) * c2[i];
    for (int i = 0; i < 3; ++i) t[i] = dir[i] + half_h * st.k2[i];
    c1 = cross(t, st.Bmiddle);
    c2 = cross(dk2dL, st.Bmiddle);
    for (int i = 0; i < 3; ++i) dk3dL[i] = c1[i] + (qop * half_h) * c2[i];
    for (int i = 0; i < 3; ++i) t[i] = dir[i] + h * st.k3[i];
    c1 = cross(t, st.Blast);
    c2 = cross(dk3dL, st.Blast);
    for (int i = 0; i < 3; ++i) dk4dL[i] = c1[i] + (qop * h) * c2[i];

    Mat<3, 3> dk1dT = Mat<3, 3>::zero();
    dk1dT.a[0][1] = st.Bfirst[2];
    dk1dT.a[0][2] = -st.Bfirst[1];
    dk1dT.a[1][0] = -st.Bfirst[2];
    dk1dT.a[1][2] = st.Bfirst[0];
    dk1dT.a[2][0] = st.Bfirst[1];
    dk1dT.a[2][1] = -st.Bfirst[0];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) dk1dT.a[i][j] *= qop;
    auto stage = [&](const Mat<3, 3>& prev, double hh, const V3& B) {
      Mat<3, 3> m3 = Mat<3, 3>::identity();
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m3.a[i][j] += hh * prev.a[i][j];
      Mat<3, 3> c = crossCols(m3, B);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) c.a[i][j] = qop * c.a[i][j];
      return c;
    };
    Mat<3, 3> dk2dT = stage(dk1dT, half_h, st.Bmiddle);
    Mat<3, 3> dk3dT = stage(dk2dT, half_h, st.Bmiddle);
    Mat<3, 3> dk4dT = stage(dk3dT, h, st.Blast);
    const double h6 = h / 6.;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double dF = (i == j ? 1.0 : 0.0) +
                    h6 * ((dk1dT.a[i][j] + dk2dT.a[i][j]) + dk3dT.a[i][j]);
        D.a[i][4 + j] = dF * h;
        D.a[4 + i][4 + j] =
            D.a[4 + i][4 + j] +
            h6 * ((dk1dT.a[i][j] + 2. * (dk2dT.a[i][j] + dk3dT.a[i][j])) +
                  dk4dT.a[i][j]);
      }
    const double hh6 = (h * h) / 6.;
    for (int i = 0; i < 3; ++i) {
      D.a[i][7] = hh6 * ((dk1dL[i] + dk2dL[i]) + dk3dL[i]);
      D.a[4 + i][7] =
          h6 * ((dk1dL[i] + 2. * (dk2dL[i] + dk3dL[i])) + dk4dL[i]);
    }
    D.a[3][7] = h * m * m / (p * p * qop * dtds);

    // EigenStepper.ipp:341-347 block update of the free transport jacobian
    Mat<4, 4> DTR, DBR, JTR, JBR;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        DTR.a[i][j] = D.a[i][4 + j];
        DBR.a[i][j] = D.a[4 + i][4 + j];
        JTR.a[i][j] = st.jacTransport.a[i][4 + j];
        JBR.a[i][j] = st.jacTransport.a[4 + i][4 + j];
      }
    Mat<4, 4> prodT = mul(DTR, JBR);
    Mat<4, 4> prodB = mul(DBR, JBR);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        st.jacTransport.a[i][4 + j] = JTR.a[i][j] + prodT.a[i][j];
        st.jacTransport.a[4 + i][4 + j] = prodB.a[i][j];
      }
  }

  // EigenStepper.ipp:355-359
  for (int i = 0; i < 3; ++i)
    st.pars.a[i][0] = st.pars.a[i][0] +
                      (h * dir[i] + (h2 / 6.) * ((st.k1[i] + st.k2[i]) + st.k3[i]));
  for (int i = 0; i < 3; ++i)
    st.pars.a[4 + i][0] =
        st.pars.a[4 + i][0] +
        (h / 6.) * ((st.k1[i] + 2. * (st.k2[i] + st.k3[i])) + st.k4[i]);
  {
    V3 d = st.direction();
    double z = squaredNorm(d);
    if (z > 0) {
      double n = std::sqrt(z);
      for (int i = 0; i < 3; ++i) st.pars.a[4 + i][0] = d[i] / n;
    }
  }
  if (st.covTransport) {
    for (int i = 0; i < 3; ++i) st.derivative.a[i][0] = st.pars.a[4 + i][0];
    for (int i = 0; i < 3; ++i) st.derivative.a[4 + i][0] = st.k4[i];
  }
  st.pathAccumulated += h;
  ++st.nSteps;
  st.nStepTrials += nTrials;

  // EigenStepper.ipp:379-385
  const double sc = scaling(errorEstimate);
  const double nextAccuracy = std::abs(h * sc);
  const double previousAccuracy = std::abs(st.stepSize.accuracy);
  const double initialStepLength = std::abs(initialH);
  if (nextAccuracy < initialStepLength || nextAccuracy > previousAccuracy)
    st.stepSize.accuracy = nextAccuracy;
  return {0, h};
}

// ===========================================================================
// CovarianceEngine.cpp / JacobianEngine.cpp
// ===========================================================================

// JacobianEngine.cpp:122-141 (the Result is ignored by the caller)
void reinitializeJacobians(const Surface& s, Stepper& st) {
  st.jacTransport = M8::identity();
  st.derivative = Vec8::zero();
  V3 pos = st.position(), dir = st.direction();
  double l0, l1;
  if (!s.globalToLocal(pos, dir, l0, l1)) return;
  st.jacToGlobal = s.boundToFreeJacobian(pos, dir);
}

// CovarianceEngine.cpp:102-165 + JacobianEngine.cpp:19-45
void transportCovarianceToBound(const Surface& s, Stepper& st) {
  V3 pos = st.position(), dir = st.direction();
  Mat<1, 8> freeToPath = s.freeToPathDerivative(pos, dir);
  Mat<6, 8> F2B = s.freeToBoundJacobian(pos, dir);
  M8 corr = M8::identity();
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j)
      corr.a[i][j] = corr.a[i][j] + st.derivative.a[i][0] * freeToPath.a[0][j];
  Mat<6, 8> A = mul(F2B, corr);
  Mat<6, 8> B = mul(A, st.jacTransport);
  st.jacobian = mul(B, st.jacToGlobal);
  M6 JC = mul(st.jacobian, st.cov);
  st.cov = mulT(JC, st.jacobian);
  reinitializeJacobians(s, st);
}

// CovarianceEngine.cpp:29-63 (returns status; params out)
int boundState(const Surface& s, Stepper& st, bool transportCov, Vec6& params) {
  int r = freeToBound(st.pars, s, params);
  if (r) return r;
  if (st.covTransport && transportCov) transportCovarianceToBound(s, st);
  return 0;
}

// EigenStepper.ipp:37-70
void initializeStepper(Stepper& st, const Surface& s, const Vec6& bp,
                       const M6* cov, const acts_kf_options& o) {
  st.pars = boundToFree(s, bp);
  st.pathAccumulated = 0;
  st.nSteps = 0;
  st.nStepTrials = 0;
  st.stepSize = ConstrainedStep();
  st.stepSize.accuracy = o.initial_step_size;
  st.stepSize.values[ConstrainedStep::User] = o.max_step_size;
  st.previousStepSize = 0;
  st.covTransport = cov != nullptr;
  if (st.covTransport) {
    st.cov = *cov;
    st.jacToGlobal = s.boundToFreeJacobian(st.position(), st.direction());
    st.jacobian = M6::identity();
    st.jacTransport = M8::identity();
    st.derivative = Vec8::zero();
  }
}

// ===========================================================================
// PointwiseMaterialInteraction.hpp/.cpp
// ===========================================================================
enum MatMode { NoUpdate = 0, PreUpdate = 1, PostUpdate = 2, FullUpdate = 3 };

// ORACLE_PLANT selects a deliberately wrong material application, compiled
// only into the test libraries liboracle_kf_plant<N>.so, so that the tests
// can show that their closed-form checks detect it:
//   1: material applied before the update on measurement surfaces (forward
//      PreUpdate factor 1, PostUpdate factor 0),
//   2: phi variance theta0^2 without the 1/sin^2(theta) factor,
//   3: no 10 MeV momentum floor after the energy loss;
// and in the layer navigator:
//   4: the layer list resolved once at the start (not after each reached
//      layer),
//   5: layers intersected without their bounds.
// The oracle proper is built with 0.
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif

int performMaterialInteraction(const Surface& s, Stepper& st,
                               const Particle& ph, const acts_kf_options& o,
                               int mode, double propDir) {
  if (!s.hasMaterial) return 0;
  // evaluateMaterialSlab (PointwiseMaterialInteraction.cpp:31-50)
  V3 pos = st.position(), dir = st.direction();
  const double pathCorrection = s.pathCorrection(pos, dir);
  double l0, l1;
  if (!s.globalToLocal(pos, l0, l1)) return ACTS_KF_STATUS_NOT_ON_SURFACE;
  SlabF slab = s.slab;
  if (!slab.isVacuum()) {
    // ISurfaceMaterial.cpp:18-46, split factor 1
    double factor = 0.;
    if (mode == FullUpdate)
      factor = 1.;
    else if (mode == PreUpdate)
      factor = (ORACLE_PLANT == 1) ? 1. : (propDir < 0 ? 1. : 0.);
    else if (mode == PostUpdate)
      factor = (ORACLE_PLANT == 1) ? 0. : (propDir > 0 ? 1. : 0.);
    if (factor == 0.) return 0;  // MaterialSlab::Nothing()
    slab.scaleThickness(static_cast<float>(factor));
  }
  slab.scaleThickness(static_cast<float>(pathCorrection));
  if (slab.isVacuum()) return 0;

  // computeMaterialEffects (PointwiseMaterialInteraction.cpp:52-87)
  const float qopF = static_cast<float>(st.qop());
  const double mass = ph.mass;
  const double absQ = ph.absQ;
  double eLoss = 0, varPhi = 0, varTheta = 0, varQop = 0;
  if (o.energy_loss)
    eLoss = energyLossBethe(slab, static_cast<float>(mass), qopF,
                            static_cast<float>(absQ));
  if (st.covTransport) {
    if (o.multiple_scattering) {
      const double theta0 = multipleScatteringTheta0(
          slab, ph.absPdg, static_cast<float>(mass), qopF, static_cast<float>(absQ));
      const double sigmaPhi = (ORACLE_PLANT == 2) ? theta0 : theta0 * (norm(dir) / perp(dir));
      varPhi = sigmaPhi * sigmaPhi;
      varTheta = theta0 * theta0;
    }
    if (o.energy_loss) {
      const double sq = energyLossLandauSigmaQOverP(
          slab, static_cast<float>(mass), qopF, static_cast<float>(absQ));
      varQop = sq * sq;
    }
  }

  // performMaterialInteraction (PointwiseMaterialInteraction.hpp:165-226)
  const double momentum = ph.extractMomentum(st.qop());
  const double nextE = std::sqrt(mass * mass + momentum * momentum) - eLoss * propDir;
  double nextP = (mass < nextE) ? fastCathetus1(nextE, mass) : 0;
  if (ORACLE_PLANT != 3) nextP = std::max(10 * kMeV, nextP);
  const double nextQop =
      ph.qOverP(nextP, static_cast<float>(std::copysign(absQ, qopF)));
  st.pars.a[7][0] = nextQop;
  st.cov.a[2][2] = std::max(0., st.cov.a[2][2] + std::copysign(varPhi, 1));
  st.cov.a[3][3] = std::max(0., st.cov.a[3][3] + std::copysign(varTheta, 1));
  st.cov.a[4][4] = std::max(0., st.cov.a[4][4] + std::copysign(varQop, 1));
  return 0;
}

}  // namespace

// ===========================================================================
// GainMatrixUpdaterImpl.hpp:21-91 (kMeasurementSize = 2, subspace {0, 1})
// ===========================================================================
// The projector H (N x 6) of a measurement's subspace
// (FixedBoundSubspaceHelper::projector)
template <int N>
Mat<N, 6> projector(const Meas& m) {
  Mat<N, 6> H = Mat<N, 6>::zero();
  for (int i = 0; i < N; ++i) H.a[i][m.idx[i]] = 1;
  return H;
}

// Eigen's closed-form inverses of a 1x1 and a 2x2 matrix
inline Mat<1, 1> inverseSmall(const Mat<1, 1>& S) {
  Mat<1, 1> r;
  r.a[0][0] = 1.0 / S.a[0][0];
  return r;
}
inline Mat<2, 2> inverseSmall(const Mat<2, 2>& S) {
  Mat<2, 2> Si;
  const double det = S.a[0][0] * S.a[1][1] - S.a[1][0] * S.a[0][1];
  const double invdet = 1.0 / det;
  Si.a[0][0] = S.a[1][1] * invdet;
  Si.a[1][0] = -S.a[1][0] * invdet;
  Si.a[0][1] = -S.a[0][1] * invdet;
  Si.a[1][1] = S.a[0][0] * invdet;
  return Si;
}

template <int N>
int gainMatrixUpdateN(const Vec6& x, const M6& P, const Meas& ms, bool useJoseph,
                      Vec6& filtered, M6& filteredCov, double& chi2) {
  const Mat<N, 6> H = projector<N>(ms);
  Mat<N, N> Vm;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) Vm.a[i][j] = ms.V[N * i + j];
  Mat<6, N> PHt = mulT(P, H);
  Mat<N, 6> HP = mul(H, P);
  Mat<N, N> S = mulT(HP, H);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) S.a[i][j] = S.a[i][j] + Vm.a[i][j];
  const Mat<N, N> Si = inverseSmall(S);
  Mat<6, N> K = mul(PHt, Si);
  if (hasNaN(K)) return ACTS_KF_STATUS_UPDATE_FAILED;

  Mat<N, 1> Hx = mul(H, x);
  Mat<N, 1> r;
  for (int i = 0; i < N; ++i) r.a[i][0] = ms.v[i] - Hx.a[i][0];
  Mat<6, 1> Kr = mul(K, r);
  for (int i = 0; i < 6; ++i) filtered.a[i][0] = x.a[i][0] + Kr.a[i][0];
  normalizePhiTheta(filtered.a[2][0], filtered.a[3][0]);

  M6 KH = mul(K, H);
  M6 tmp;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) tmp.a[i][j] = (i == j ? 1.0 : 0.0) - KH.a[i][j];
  if (!useJoseph) {
    filteredCov = mul(tmp, P);
  } else {
    M6 a = mulT(mul(tmp, P), tmp);
    Mat<6, N> KV = mul(K, Vm);
    M6 b = mulT(KV, K);
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) filteredCov.a[i][j] = a.a[i][j] + b.a[i][j];
  }

  Mat<N, 1> Hf = mul(H, filtered);
  Mat<N, 1> res;
  for (int i = 0; i < N; ++i) res.a[i][0] = ms.v[i] - Hf.a[i][0];
  Mat<N, N> HK = mul(H, K);
  Mat<N, N> IHK;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) IHK.a[i][j] = (i == j ? 1.0 : 0.0) - HK.a[i][j];
  const Mat<N, N> mi = inverseSmall(mul(IHK, Vm));
  // (r^T * m^-1) * r
  chi2 = 0;
  for (int j = 0; j < N; ++j) {
    double row = res.a[0][0] * mi.a[0][j];
    for (int i = 1; i < N; ++i) row = row + res.a[i][0] * mi.a[i][j];
    chi2 = (j == 0) ? row * res.a[0][0] : chi2 + row * res.a[j][0];
  }
  return 0;
}

int gainMatrixUpdate(const Vec6& x, const M6& P, const Meas& m, bool useJoseph,
                     Vec6& filtered, M6& filteredCov, double& chi2) {
  if (m.dim == 1) return gainMatrixUpdateN<1>(x, P, m, useJoseph, filtered, filteredCov, chi2);
  return gainMatrixUpdateN<2>(x, P, m, useJoseph, filtered, filteredCov, chi2);
}

int gainMatrixUpdate(const Vec6& x, const M6& P, const double meas[2],
                     const double V[4], bool useJoseph, Vec6& filtered,
                     M6& filteredCov, double& chi2) {
  Meas m;
  for (int k = 0; k < 2; ++k) m.v[k] = meas[k];
  for (int k = 0; k < 4; ++k) m.V[k] = V[k];
  return gainMatrixUpdateN<2>(x, P, m, useJoseph, filtered, filteredCov, chi2);
}

// TrackHelpers.hpp:676-694 calculatePredictedChi2 (residual of the
// prediction, its covariance H P H^T + V)
static double predictedChi2(const Vec6& x, const M6& P, const Meas& m) {
  if (m.dim == 1) {
    const double r0 = m.v[0] - x.a[m.idx[0]][0];
    const double S00 = P.a[m.idx[0]][m.idx[0]] + m.V[0];
    const double i00 = 1.0 / S00;
    return (r0 * i00) * r0;
  }
  const int a = m.idx[0], b = m.idx[1];
  double r0 = m.v[0] - x.a[a][0];
  double r1 = m.v[1] - x.a[b][0];
  double S00 = P.a[a][a] + m.V[0], S01 = P.a[a][b] + m.V[1];
  double S10 = P.a[b][a] + m.V[2], S11 = P.a[b][b] + m.V[3];
  const double det = S00 * S11 - S10 * S01;
  const double invdet = 1.0 / det;
  double i00 = S11 * invdet, i10 = -S10 * invdet, i01 = -S01 * invdet,
         i11 = S00 * invdet;
  double row0 = r0 * i00 + r1 * i10;
  double row1 = r0 * i01 + r1 * i11;
  return row0 * r0 + row1 * r1;
}

// ===========================================================================
// 6x6 inverse: partial-pivot LU (Eigen PartialPivLU::inverse for 6x6;
// last-bit agreement with Eigen's blocked triangular solves is unpinned)
// ===========================================================================
M6 inverse6(const M6& A) {
  double lu[6][6];
  int perm[6];
  for (int i = 0; i < 6; ++i) {
    perm[i] = i;
    for (int j = 0; j < 6; ++j) lu[i][j] = A.a[i][j];
  }
  for (int k = 0; k < 6; ++k) {
    int piv = k;
    double big = std::abs(lu[k][k]);
    for (int i = k + 1; i < 6; ++i)
      if (std::abs(lu[i][k]) > big) {
        big = std::abs(lu[i][k]);
        piv = i;
      }
    if (big != 0.0) {
      if (piv != k) {
        for (int j = 0; j < 6; ++j) std::swap(lu[k][j], lu[piv][j]);
        std::swap(perm[k], perm[piv]);
      }
      for (int i = k + 1; i < 6; ++i) lu[i][k] = lu[i][k] / lu[k][k];
    }
    for (int i = k + 1; i < 6; ++i)
      for (int j = k + 1; j < 6; ++j) lu[i][j] = lu[i][j] - lu[i][k] * lu[k][j];
  }
  M6 X;
  for (int c = 0; c < 6; ++c) {
    double b[6];
    for (int i = 0; i < 6; ++i) b[i] = (perm[i] == c) ? 1.0 : 0.0;
    for (int k = 0; k < 6; ++k)
      for (int i = k + 1; i < 6; ++i) b[i] = b[i] - lu[i][k] * b[k];
    for (int k = 5; k >= 0; --k) {
      b[k] = b[k] / lu[k][k];
      for (int i = 0; i < k; ++i) b[i] = b[i] - lu[i][k] * b[k];
    }
    for (int i = 0; i < 6; ++i) X.a[i][c] = b[i];
  }
  return X;
}

// GainMatrixSmoother.cpp:18-97
int gainMatrixSmoothStep(const Vec6& filtered, const M6& Cf,
                         const Vec6& prevPredicted, const M6& Cp,
                         const Vec6& prevSmoothed, const M6& Cs,
                         const M6& J, Vec6& smoothed, M6& smoothedCov) {
  M6 CfJt = mulT(Cf, J);
  M6 G = mul(CfJt, inverse6(Cp));
  if (hasNaN(G)) return ACTS_KF_STATUS_SMOOTH_FAILED;
  Vec6 d;
  for (int i = 0; i < 6; ++i) d.a[i][0] = prevSmoothed.a[i][0] - prevPredicted.a[i][0];
  d.a[2][0] = radianSym(d.a[2][0]);
  d.a[3][0] = radianSym(d.a[3][0]);
  Vec6 Gd = mul(G, d);
  for (int i = 0; i < 6; ++i) smoothed.a[i][0] = filtered.a[i][0] + Gd.a[i][0];
  normalizePhiTheta(smoothed.a[2][0], smoothed.a[3][0]);
  M6 diff;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) diff.a[i][j] = Cs.a[i][j] - Cp.a[i][j];
  M6 GdG = mulT(mul(G, diff), G);
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) smoothedCov.a[i][j] = Cf.a[i][j] + GdG.a[i][j];
  return 0;
}

// ===========================================================================
// Propagator<EigenStepper, DirectNavigator> + KalmanFitter::Actor
// ===========================================================================
namespace {

struct DirectNav {
  const std::vector<int>* seq = nullptr;
  int surfaceIndex = -1;
  int current = -1;  // -1 none, -2 start surface, >= 0 geometry index
  bool navigationBreak = false;
  double direction = 1.0;  // propagation direction
};

struct Target {
  bool none = true;
  int surface = -1;
  int index = 0;
};

// The bounds check of Surface::intersect with BoundaryTolerance::None at
// the point x: CylinderSurface.cpp:242-257 (full azimuth: |local z| below
// the half length plus the tolerance), DiscSurface.cpp:286-293 (full disc:
// squared distance from the centre within rmin^2 .. rmax^2), and for a plane
// its RectangleBounds (half lengths b0, b1; none when both are 0).
bool insideBoundsNone(const Surface& s, const V3& x, double tol) {
  const V3 rel = v3(x[0] - s.pl.t[0], x[1] - s.pl.t[1], x[2] - s.pl.t[2]);
  if (s.type == ACTS_KF_SURFACE_CYLINDER) {
    const double cZ = dot(rel, s.pl.col(2));
    return std::abs(cZ) < std::abs(s.b1 + tol);
  }
  if (s.type == ACTS_KF_SURFACE_DISC) {
    const double r2 = dot(rel, rel);
    return r2 >= s.b0 * s.b0 && r2 <= s.b1 * s.b1;
  }
  if (s.type == ACTS_KF_SURFACE_PLANE && s.b0 > 0 && s.b1 > 0) {
    const double lx = dot(rel, s.pl.col(0)), ly = dot(rel, s.pl.col(1));
    return -s.b0 <= lx && lx <= s.b0 && -s.b1 <= ly && ly <= s.b1;
  }
  return true;
}

// Surface::intersect with BoundaryTolerance::None: a solution outside the
// bounds is unreachable.
MultiIntersection intersectBounded(const Surface& s, const V3& pos, const V3& dir,
                                   double tol) {
  MultiIntersection mi = s.intersect(pos, dir, tol);
  if (ORACLE_PLANT == 5) return mi;
  for (int i = 0; i < mi.size; ++i) {
    Intersection& x = mi.is[i];
    if (!x.valid()) continue;
    const V3 at = v3(pos[0] + x.path * dir[0], pos[1] + x.path * dir[1], pos[2] + x.path * dir[2]);
    if (!insideBoundsNone(s, at, tol)) x.status = IStatus::unreachable;
  }
  return mi;
}

// SteppingHelper.hpp:40-78: BoundaryTolerance::Infinite for the direct
// navigator's targets, None for the layer navigator's (bounded)
IStatus updateSurfaceStatus(const Surface& s, Stepper& st, int index,
                            double propDir, double tol, bool bounded = false) {
  V3 d = st.direction();
  V3 dd = v3(propDir * d[0], propDir * d[1], propDir * d[2]);
  MultiIntersection mi =
      bounded ? intersectBounded(s, st.position(), dd, tol) : s.intersect(st.position(), dd, tol);
  const Intersection& x = mi.is[index];
  if (x.status == IStatus::onSurface) {
    st.stepSize.release(ConstrainedStep::Navigator);
    st.updateStepSize(x.path, ConstrainedStep::Navigator);
    return IStatus::onSurface;
  }
  const double nearLimit = std::numeric_limits<double>::lowest();
  const double farLimit = std::numeric_limits<double>::max();
  if (x.valid() && x.path > nearLimit && x.path < farLimit + kOnSurfaceTolerance) {
    st.stepSize.release(ConstrainedStep::Navigator);
    st.updateStepSize(x.path, ConstrainedStep::Navigator);
    return IStatus::reachable;
  }
  return IStatus::unreachable;
}

// DirectNavigator.hpp:280-321, 371-390
Target nextTarget(const Geometry& geo, DirectNav& nav, const V3& pos,
                  const V3& dir) {
  nav.current = -1;
  if (nav.navigationBreak) return Target{};
  // nextSurface / endOfSurfaces (DirectNavigator.hpp:107-124): the sequence
  // in the propagation direction
  const int n = static_cast<int>(nav.seq->size());
  const int inc = nav.direction < 0 ? -1 : 1;
  nav.surfaceIndex += inc;
  const V3 d = v3(nav.direction * dir[0], nav.direction * dir[1], nav.direction * dir[2]);
  while (nav.surfaceIndex >= 0 && nav.surfaceIndex < n) {
    const int sid = (*nav.seq)[nav.surfaceIndex];
    const Surface& s = geo.surfaces[sid];
    MultiIntersection mi = s.intersect(pos, d, geo.opt.surface_tolerance);
    for (int i = 0; i < mi.size; ++i) {
      const Intersection& x = mi.is[i];
      if (x.valid() && x.path > geo.opt.near_limit &&
          x.path < std::numeric_limits<double>::max() + kOnSurfaceTolerance) {
        return Target{false, sid, i};
      }
    }
    nav.surfaceIndex += inc;
  }
  nav.navigationBreak = true;
  return Target{};
}

// The layer navigator: Navigator with Gen1 targeting (Navigator.cpp:
// 251-316 nextTarget and renavigation, :479-540 getNextTargetGen1, :826-849
// resolveLayers) over the single surfaces of the geometry table as layers,
// resolved as the reference's fitter tests configure it (resolveSensitive
// and resolveMaterial on, resolvePassive off, FitterTestsCommon.hpp:96-99).
// Such a layer has no surface array, so reaching it yields no further
// surfaces (Layer::compatibleSurfaces, Layer.cpp:116-126).  Each reached
// layer is treated like a volume switch (resetAfterVolumeSwitch,
// Navigator.hpp:245-258): the remaining layers are resolved again from it,
// with it as the start object.  (Gen1 resolves a volume's list once; with
// cylinders and discs crossing each other in one volume, as in the endcap
// detectors here, a list resolved along the start direction misorders
// nearly coincident crossings.)  As in a Gen1 list, which holds each layer
// of a volume once, a layer that has been reached (or the start surface) is
// not resolved again: where a cylinder and a disc cross each other the
// track would otherwise go back and forth between the two.  A used-up list
// means the boundaries, of which there are none, so the navigator
// renavigates.
struct LayerNav {
  struct Cand {
    double path;
    int surface;
    int index;
  };
  std::vector<Cand> list;  // resolved layers, by path length (ties: index)
  size_t next = 0;         // the next entry to target
  bool resolved = false;
  std::vector<char> reached;  // layers reached (and the start surface)
  int current = -1;        // -1 none, -2 start surface, >= 0 geometry index
  bool navigationBreak = false;
  double direction = 1.0;
};

// TrackingVolume::compatibleLayers (TrackingVolume.cpp:443-487) with
// Layer::resolve and Layer::surfaceOnApproach (Layer.hpp:147-168,
// Layer.cpp:224-266: no approach descriptor, so the layer surface's first
// intersection inside its bounds within the path window), sorted as
// resolveLayers sorts them.  The near limit is NavigatorPlainOptions'
// (the on-surface tolerance, NavigatorOptions.hpp:42), the far limit
// unbounded; the layers already reached are left out (the start object).
// Here the layers are not walked through a binned layer array (nextLayer),
// so the list is sorted by path length from the start and TrackingVolume's
// rotation to the closest entry changes nothing.
void resolveLayers(const Geometry& geo, LayerNav& nav, const V3& pos, const V3& dir) {
  nav.list.clear();
  nav.next = 0;
  const int n = static_cast<int>(geo.surfaces.size());
  for (int sid = 0; sid < n; ++sid) {
    if (nav.reached[sid]) continue;
    const Surface& s = geo.surfaces[sid];
    if (!s.sensitive && !s.hasMaterial) continue;
    const MultiIntersection mi = intersectBounded(s, pos, dir, kOnSurfaceTolerance);
    for (int i = 0; i < mi.size; ++i) {
      const Intersection& x = mi.is[i];
      if (x.valid() && x.path > kOnSurfaceTolerance &&
          x.path < std::numeric_limits<double>::max() + kOnSurfaceTolerance) {
        nav.list.push_back({x.path, sid, i});
        break;
      }
    }
  }
  std::sort(nav.list.begin(), nav.list.end(), [](const LayerNav::Cand& a, const LayerNav::Cand& b) {
    return a.path < b.path || (a.path == b.path && a.surface < b.surface);
  });
}

// Navigator::nextTarget for the layer navigator: after a reached layer (or
// at the start) the list is resolved from here; else the next layer of the
// list; a used-up list is resolved again from the current position
// (renavigation); if that finds none, the navigation ends.
Target nextTargetLayers(const Geometry& geo, LayerNav& nav, const V3& pos, const V3& dir) {
  const int on = nav.current;
  nav.current = -1;
  if (nav.navigationBreak) return Target{};
  if (on >= 0) nav.reached[on] = 1;
  const V3 d = v3(nav.direction * dir[0], nav.direction * dir[1], nav.direction * dir[2]);
  if (!nav.resolved || (on >= 0 && ORACLE_PLANT != 4)) {
    resolveLayers(geo, nav, pos, d);
    nav.resolved = true;
  }
  if (nav.next < nav.list.size()) {
    const LayerNav::Cand& c = nav.list[nav.next++];
    return Target{false, c.surface, c.index};
  }
  resolveLayers(geo, nav, pos, d);
  if (nav.next < nav.list.size()) {
    const LayerNav::Cand& c = nav.list[nav.next++];
    return Target{false, c.surface, c.index};
  }
  nav.navigationBreak = true;
  return Target{};
}

// LoopProtection.hpp:19-67
void setupLoopProtection(const Geometry& geo, const acts_kf_options& o, Stepper& st,
                         double& limit, bool releaseLimit, double propDir) {
  if (!o.loop_protection) return;
  if (releaseLimit) limit = propDir * std::numeric_limits<double>::max();
  V3 B;
  if (geo.field.get(st.position(), B)) return;
  const double Bn = norm(B);
  if (Bn == 0) return;
  const double p = geo.particle.extractMomentum(st.qop());
  const double helixPath = propDir * 2 * kPi * p / Bn;
  const double loopLimit = o.loop_fraction * helixPath;
  if (std::abs(loopLimit) < std::abs(limit)) limit = loopLimit;
}
void setupLoopProtection(const Geometry& geo, Stepper& st, double& limit,
                         bool releaseLimit, double propDir) {
  setupLoopProtection(geo, geo.opt, st, limit, releaseLimit, propDir);
}

// StandardAborters.hpp:41-61
bool pathLimitCheckAbort(double internalLimit, Stepper& st, double tol) {
  double distance = std::abs(internalLimit) - std::abs(st.pathAccumulated);
  if (std::abs(distance) < std::abs(tol)) return true;
  st.updateStepSize(distance, ConstrainedStep::Actor);
  return false;
}

struct KFState {
  StateOut s;
  bool isMeasurementBranch = false;
};

struct KFResult {
  std::vector<KFState> states;
  int lastMeasurementIndex = -1;
  int lastTrackIndex = -1;
  size_t measurementStates = 0;
  size_t processedStates = 0;
  size_t measurementHoles = 0;
  size_t missedActive = 0;
  bool finished = false;
  double pathLimit = std::numeric_limits<double>::max();
};

struct MeasLookup {
  std::vector<int> surfaces;  // unique surfaces in insertion order
  std::vector<int> index;     // measurement index for each surface
  int find(int sid) const {
    for (size_t i = 0; i < surfaces.size(); ++i)
      if (surfaces[i] == sid) return index[i];
    return -1;
  }
};

// KalmanFitter.hpp:445-671 for a surface of the sequence
int kfFilter(const Geometry& geo, const TrackIn& in, const MeasLookup& ml,
             const Surface& surf, int sid, bool isStartSurface, Stepper& st,
             KFResult& res, double propDir) {
  const bool precedingMeasurementExists = res.measurementStates > 0;
  const bool sensitive = surf.sensitive;
  const bool hasMat = surf.hasMaterial;
  // the track's own start plane (sid -2) carries no measurement; a geometry
  // surface does when a measurement is keyed to it (the reverse filter
  // starts on the last measurement surface)
  const int mi = sid >= 0 ? ml.find(sid) : -1;
  const auto& o = geo.opt;
  // determineMaterialUpdateMode (PointwiseMaterialInteraction.cpp:16-29):
  // the start surface only takes post-update material
  auto matMode = [&](int requested) {
    if (isStartSurface) return requested & PostUpdate;
    return requested;
  };
  if (mi >= 0) {
    transportCovarianceToBound(surf, st);
    int r = performMaterialInteraction(surf, st, geo.particle, o,
                                       matMode(PreUpdate), propDir);
    if (r) return r;
    KFState ks;
    ks.isMeasurementBranch = true;
    ks.s.meas = mi;
    Vec6 bp;
    r = boundState(surf, st, false, bp);
    if (r) return r;
    // the track's state slots (only the layer navigator can find more)
    if (propDir > 0 && res.states.size() >= in.sequence.size())
      return ACTS_KF_STATUS_STATE_CAPACITY;
    ks.s.surface = sid;
    ks.s.predicted = bp;
    ks.s.predictedCov = st.cov;
    ks.s.jacobian = st.jacobian;
    ks.s.path = st.pathAccumulated;
    ks.s.flags = ACTS_KF_STATE_HAS_PARAMETERS;
    if (hasMat) ks.s.flags |= ACTS_KF_STATE_HAS_MATERIAL;
    const Meas ms = in.meas(mi);
    bool outlier = false;
    if (o.outlier_chi2_cut > 0) {
      outlier = predictedChi2(bp, st.cov, ms) > o.outlier_chi2_cut;
    }
    if (!outlier) {
      double chi2 = 0;
      r = gainMatrixUpdate(ks.s.predicted, ks.s.predictedCov, ms,
                           o.use_joseph != 0, ks.s.filtered, ks.s.filteredCov,
                           chi2);
      if (r) return r;
      ks.s.chi2 = static_cast<float>(chi2);
      ks.s.flags |= ACTS_KF_STATE_HAS_MEASUREMENT;
    } else {
      ks.s.flags |= ACTS_KF_STATE_HAS_MEASUREMENT | ACTS_KF_STATE_IS_OUTLIER;
      ks.s.filtered = ks.s.predicted;
      ks.s.filteredCov = ks.s.predictedCov;
    }
    res.states.push_back(ks);
    res.lastTrackIndex = static_cast<int>(res.states.size()) - 1;
    if (!outlier) {
      // EigenStepper::update (EigenStepper.ipp:129-138)
      st.pars = boundToFree(surf, ks.s.filtered);
      st.cov = ks.s.filteredCov;
      st.jacToGlobal = surf.boundToFreeJacobian(st.position(), st.direction());
      ++res.measurementStates;
    }
    r = performMaterialInteraction(surf, st, geo.particle, o, matMode(PostUpdate),
                                   propDir);
    if (r) return r;
    ++res.processedStates;
    res.measurementHoles = res.missedActive;
    res.lastMeasurementIndex = res.lastTrackIndex;
  } else if ((precedingMeasurementExists && sensitive) || hasMat) {
    KFState ks;
    Vec6 bp;
    int r = boundState(surf, st, true, bp);
    if (r) return r;
    if (propDir > 0 && res.states.size() >= in.sequence.size())
      return ACTS_KF_STATUS_STATE_CAPACITY;
    ks.s.surface = sid;
    ks.s.predicted = bp;
    ks.s.predictedCov = st.cov;
    ks.s.jacobian = st.jacobian;
    ks.s.path = st.pathAccumulated;
    ks.s.filtered = bp;
    ks.s.filteredCov = st.cov;
    ks.s.flags = ACTS_KF_STATE_HAS_PARAMETERS;
    if (hasMat) ks.s.flags |= ACTS_KF_STATE_HAS_MATERIAL;
    if (sensitive && precedingMeasurementExists) ks.s.flags |= ACTS_KF_STATE_IS_HOLE;
    res.states.push_back(ks);
    res.lastTrackIndex = static_cast<int>(res.states.size()) - 1;
    if (ks.s.flags & ACTS_KF_STATE_IS_HOLE) ++res.missedActive;
    ++res.processedStates;
    r = performMaterialInteraction(surf, st, geo.particle, o, matMode(FullUpdate),
                                   propDir);
    if (r) return r;
  }
  return 0;
}

}  // namespace

// MbfSmoother::operator() (MbfSmoother.hpp:48-89, MbfSmoother.cpp:16-48,
// detail/MbfSmootherImpl.hpp:17-57): from the tip backwards, smoothed =
// filtered - C_f lambda, C_s = C_f - C_f Lambda C_f; then Lambda, lambda
// pass through the state's measurement update (2D, projector on loc0/loc1)
// if it is a measurement, and through its transport jacobian F.  Products
// are associated as Eigen evaluates them (left to right).
void mbfSmooth(const TrackIn& in, std::vector<KFState>& S, int tip) {
  M6 bigL = M6::zero();
  Vec6 smallL = Vec6::zero();
  for (int k = tip; k >= 0; --k) {
    StateOut& s = S[k].s;
    // calculateSmoothed
    const M6 CfLCf = mul(mul(s.filteredCov, bigL), s.filteredCov);
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) s.smoothedCov.a[i][j] = s.filteredCov.a[i][j] - CfLCf.a[i][j];
    const Vec6 Cfl = mul(s.filteredCov, smallL);
    for (int i = 0; i < 6; ++i) s.smoothed.a[i][0] = s.filtered.a[i][0] - Cfl.a[i][0];
    normalizePhiTheta(s.smoothed.a[2][0], s.smoothed.a[3][0]);
    if (k == 0) break;
    const M6& F = s.jacobian;
    const M6 Ft = transpose(F);
    const bool isMeasurement =
        (s.flags & ACTS_KF_STATE_HAS_MEASUREMENT) && !(s.flags & ACTS_KF_STATE_IS_OUTLIER);
    if (isMeasurement) {
      const M6& P = s.predictedCov;
      const Meas ms = in.meas(s.meas);
      const int N = ms.dim;
      const int* ix = ms.idx;
      // S = H P H^T + V and its inverse (Eigen's closed-form 1x1 / 2x2
      // inverse)
      double Si[2][2] = {{0, 0}, {0, 0}};
      if (N == 1) {
        Si[0][0] = 1.0 / (P.a[ix[0]][ix[0]] + ms.V[0]);
      } else {
        const double S00 = P.a[ix[0]][ix[0]] + ms.V[0], S01 = P.a[ix[0]][ix[1]] + ms.V[1];
        const double S10 = P.a[ix[1]][ix[0]] + ms.V[2], S11 = P.a[ix[1]][ix[1]] + ms.V[3];
        const double det = S00 * S11 - S10 * S01;
        const double invdet = 1.0 / det;
        Si[0][0] = S11 * invdet;
        Si[0][1] = -S01 * invdet;
        Si[1][0] = -S10 * invdet;
        Si[1][1] = S00 * invdet;
      }
      // K = (P H^T) S^-1, CHat = I - K H, y = m - H x_pred
      double K[6][2];
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < N; ++j) {
          double acc = P.a[i][ix[0]] * Si[0][j];
          if (N == 2) acc = acc + P.a[i][ix[1]] * Si[1][j];
          K[i][j] = acc;
        }
      M6 CHat = M6::identity();
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < N; ++j) CHat.a[i][ix[j]] = CHat.a[i][ix[j]] - K[i][j];
      double y[2] = {0, 0};
      for (int j = 0; j < N; ++j) y[j] = ms.v[j] - s.predicted.a[ix[j]][0];
      // Lambda~ = (H^T S^-1) H + (CHat^T Lambda) CHat
      M6 bigLT = mul(mul(transpose(CHat), bigL), CHat);
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) bigLT.a[ix[i]][ix[j]] = Si[i][j] + bigLT.a[ix[i]][ix[j]];
      // lambda~ = (-H^T S^-1) y + CHat^T lambda
      Vec6 smallLT = mul(transpose(CHat), smallL);
      for (int i = 0; i < N; ++i) {
        double acc = -Si[i][0] * y[0];
        if (N == 2) acc = acc + -Si[i][1] * y[1];
        smallLT.a[ix[i]][0] = acc + smallLT.a[ix[i]][0];
      }
      bigL = mul(mul(Ft, bigLT), F);
      smallL = mul(Ft, smallLT);
    } else {
      bigL = mul(mul(Ft, bigL), F);
      smallL = mul(Ft, smallL);
    }
  }
}

namespace {

// What one filter pass leaves: the actor's result, the propagation's
// status and statistics, and, if the actor's target surface was reached,
// the bound parameters there (KalmanFitterResult::fittedParameters).
struct PassOut {
  KFResult res;
  int status = 0;
  int nSteps = 0, nTrials = 0;
  bool targetReached = false;
  Vec6 targetParams;
  M6 targetCov;
};

// SurfaceReached::checkAbort (StandardAborters.hpp:94-155) for the actor's
// target: reached if the closest intersection is on the surface; the
// actor step size is set to the first intersection in front of the near
// limit (updateSingleSurfaceStatus, SteppingHelper.hpp:41-85, which
// releases the actor constraint first).
bool targetReachedCheck(const Surface& target, Stepper& st, double propDir, double tol) {
  const V3 d = st.direction();
  const V3 dd = v3(propDir * d[0], propDir * d[1], propDir * d[2]);
  const MultiIntersection mi = target.intersect(st.position(), dd, tol);
  // closest(): the target is a line surface here, with one intersection
  const bool reached = mi.size > 0 && mi.is[0].status == IStatus::onSurface;
  const double nearLimit = -100 * 1e-3;
  for (int i = 0; i < mi.size; ++i) {
    const Intersection& x = mi.is[i];
    if (x.valid() && x.path > nearLimit &&
        x.path < std::numeric_limits<double>::max() + kOnSurfaceTolerance) {
      st.stepSize.release(ConstrainedStep::Actor);
      st.updateStepSize(x.path, ConstrainedStep::Actor);
      break;
    }
  }
  return reached;
}

// filter_impl (KalmanFitter.hpp:797-852): Propagator::propagate
// (Propagator.ipp:25-210) with the KalmanFitter actor, in direction
// propDir, from bound parameters on `startSurface` with covariance
// `startCov`.  startSid is the start surface's geometry index, or -2 for
// the track's own start plane, which the sequence does not contain.  The
// DirectNavigator starts next to the start surface's first entry in the
// sequence, or at the sequence's end in the propagation direction
// (DirectNavigator.hpp:222-272); the layer navigator resolves its first
// list without the start surface; the actor acts on the start surface first
// (the navigator's current surface).  `target` is the actor's targetReached
// surface (null: none).
void filterPass(const Geometry& geo, const TrackIn& in, const MeasLookup& ml,
                size_t nInputMeas, const Surface& startSurface, int startSid,
                const Vec6& startParams, const M6& startCov, double propDir,
                const Surface* target, PassOut& P) {
  const auto& o = geo.opt;
  Stepper st;
  initializeStepper(st, startSurface, startParams, &startCov, o);
  const bool layers = o.navigator == ACTS_KF_NAVIGATOR_LAYERS;
  LayerNav lnav;
  lnav.direction = propDir;
  lnav.current = startSid;
  lnav.reached.assign(geo.surfaces.size(), 0);
  if (startSid >= 0) lnav.reached[startSid] = 1;
  DirectNav nav;
  nav.seq = &in.sequence;
  nav.direction = propDir;
  nav.current = startSid;
  int& current = layers ? lnav.current : nav.current;
  const bool& navigationBreak = layers ? lnav.navigationBreak : nav.navigationBreak;
  const int n = static_cast<int>(in.sequence.size());
  int found = -1;
  for (int i = 0; i < n && startSid >= 0; ++i)
    if (in.sequence[i] == startSid) {
      found = i;
      break;
    }
  if (found >= 0)
    nav.surfaceIndex = found + (propDir < 0 ? 1 : -1);
  else
    nav.surfaceIndex = propDir < 0 ? n : -1;
  nav.navigationBreak = false;
  double propPathLimit = o.path_limit;
  setupLoopProtection(geo, st, propPathLimit, false, propDir);

  KFResult& res = P.res;
  int status = 0;

  // KalmanFitter::Actor::act (KalmanFitter.hpp:339-422)
  auto act = [&]() -> int {
    if (res.finished) return 0;
    if (res.pathLimit == std::numeric_limits<double>::max())
      setupLoopProtection(geo, st, res.pathLimit, true, propDir);
    if (current != -1) {
      const bool isStart = current == startSid;
      const Surface& s = current == -2 ? in.startSurface : geo.surfaces[current];
      int r = kfFilter(geo, in, ml, s, current, isStart, st, res, propDir);
      if (r) return r;
    }
    const bool complete = res.measurementStates == nInputMeas;
    const bool pathReached = pathLimitCheckAbort(res.pathLimit, st, o.surface_tolerance);
    const bool targetHit =
        target != nullptr && targetReachedCheck(*target, st, propDir, o.surface_tolerance);
    if (complete || pathReached || targetHit) {
      if (targetHit) {
        // boundState on the target: the fitted parameters there
        int r = boundState(*target, st, true, P.targetParams);
        if (r) return r;
        P.targetCov = st.cov;
        P.targetReached = true;
      }
      res.finished = true;
    }
    return 0;
  };
  auto checkAbort = [&]() -> bool {
    if (res.finished) return true;
    return pathLimitCheckAbort(propPathLimit, st, o.surface_tolerance);
  };

  // Propagator::propagate (Propagator.ipp:25-210)
  auto getNextTarget = [&](Target& t) -> int {
    for (unsigned i = 0; i < o.max_target_skipping; ++i) {
      Target nt = layers ? nextTargetLayers(geo, lnav, st.position(), st.direction())
                         : nextTarget(geo, nav, st.position(), st.direction());
      if (nt.none) {
        t = nt;
        return 0;
      }
      IStatus ps = updateSurfaceStatus(geo.surfaces[nt.surface], st, nt.index,
                                       propDir, o.surface_tolerance, layers);
      if (ps == IStatus::onSurface) continue;
      if (ps == IStatus::reachable) {
        t = nt;
        return 0;
      }
    }
    return ACTS_KF_STATUS_NEXT_TARGET_LIMIT;
  };

  status = act();
  bool terminatedNormally = false;
  if (!status && checkAbort()) {
    terminatedNormally = true;
  } else if (!status) {
    Target target;
    status = getNextTarget(target);
    unsigned steps = 0;
    for (; !status && steps < o.max_steps; ++steps) {
      StepResult sr = step(st, geo.field, geo.particle, o, propDir);
      if (sr.status) {
        status = sr.status;
        break;
      }
      st.stepSize.release(ConstrainedStep::Navigator);
      st.stepSize.release(ConstrainedStep::Actor);
      if (!target.none) {
        IStatus ps = updateSurfaceStatus(geo.surfaces[target.surface], st,
                                         target.index, propDir, o.surface_tolerance, layers);
        if (ps == IStatus::onSurface && !navigationBreak) current = target.surface;
        if (ps != IStatus::reachable) target = Target{};
      }
      status = act();
      if (status) break;
      if (checkAbort()) {
        terminatedNormally = true;
        break;
      }
      if (target.none) {
        st.stepSize.release(ConstrainedStep::Navigator);
        status = getNextTarget(target);
        if (status) break;
      }
    }
    if (!status && !terminatedNormally) status = ACTS_KF_STATUS_STEP_COUNT_LIMIT;
  }
  P.status = status;
  P.nSteps = st.nSteps;
  P.nTrials = st.nStepTrials;
}

// findFirstMeasurementState / findLastMeasurementState (TrackHelpers.hpp:
// 64-113): states with a measurement that is not an outlier, up to the tip
void measurementStateRange(const std::vector<KFState>& S, int tip, int& first, int& last) {
  first = last = -1;
  for (int k = 0; k <= tip; ++k) {
    const uint32_t f = S[k].s.flags;
    if ((f & ACTS_KF_STATE_HAS_MEASUREMENT) && !(f & ACTS_KF_STATE_IS_OUTLIER)) {
      if (first < 0) first = k;
      last = k;
    }
  }
}

}  // namespace

void fitTrack(const Geometry& geo, const TrackIn& in, TrackOut& out) {
  const auto& o = geo.opt;
  out = TrackOut();
  // make_propagator_options: measurements keyed by surface, first wins
  // (KalmanFitter.hpp:750-755)
  MeasLookup ml;
  for (size_t i = 0; i < in.measSurface.size(); ++i) {
    if (ml.find(in.measSurface[i]) < 0) {
      ml.surfaces.push_back(in.measSurface[i]);
      ml.index.push_back(static_cast<int>(i));
    }
  }
  const size_t nInputMeas = ml.surfaces.size();

  // the forward filter: from the start parameters, no target surface
  PassOut F;
  filterPass(geo, in, ml, nInputMeas, in.startSurface, -2, in.startParams, in.startCov, 1.0,
             nullptr, F);
  KFResult& res = F.res;
  const int status = F.status;
  out.nSteps = F.nSteps;
  out.nTrials = F.nTrials;
  if (status) {
    out.status = status;
    return;
  }
  if (res.measurementStates == 0) {
    out.status = ACTS_KF_STATUS_NO_MEASUREMENT_FOUND;
    return;
  }

  // states of the track: first .. tip (= lastMeasurementIndex)
  const int tip = res.lastMeasurementIndex;
  std::vector<KFState>& S = res.states;
  int firstMeas = -1, lastMeas = -1;
  measurementStateRange(S, tip, firstMeas, lastMeas);

  // fit_impl (KalmanFitter.hpp:884-945): reverse filtering when asked for,
  // or when the reverse-filtering logic accepts the last measurement state
  // (the Examples' SimpleReverseFilteringLogic, KalmanFitterFunction.cpp:
  // 53-57: |1 / filtered q/p| <= momentum threshold)
  const bool doReverse =
      o.reverse_filtering != 0 ||
      std::abs(1.0 / S[lastMeas].s.filtered.a[4][0]) <= o.reverse_filtering_momentum_threshold;
  out.reversed = doReverse;
  if (doReverse) {
    // the last measurement state's filtered parameters, covariance scaled
    const KFState& L = S[lastMeas];
    M6 C = L.s.filteredCov;
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) C.a[i][j] = C.a[i][j] * o.reverse_filtering_cov_scaling;
    PassOut R;
    const Surface* target =
        o.reference_surface == ACTS_KF_REFERENCE_PERIGEE ? &geo.reference : nullptr;
    filterPass(geo, in, ml, nInputMeas, geo.surfaces[L.s.surface], L.s.surface, L.s.filtered, C,
               -1.0, target, R);
    if (!R.status && R.res.measurementStates == 0) R.status = ACTS_KF_STATUS_NO_MEASUREMENT_FOUND;
    if (R.status) {
      out.status = R.status;
      return;
    }
    int rFirst = -1, rLast = -1;
    measurementStateRange(R.res.states, R.res.lastMeasurementIndex, rFirst, rLast);
    if (R.res.states[rLast].s.surface != S[firstMeas].s.surface) {
      out.status = ACTS_KF_STATUS_INCONSISTENT_TRACK_STATES;
      return;
    }
    // only the first and last measurement states carry smoothed parameters
    const double nan = std::numeric_limits<double>::quiet_NaN();
    for (int k = 0; k <= tip; ++k) {
      for (int i = 0; i < 6; ++i) {
        S[k].s.smoothed.a[i][0] = nan;
        for (int j = 0; j < 6; ++j) S[k].s.smoothedCov.a[i][j] = nan;
      }
    }
    S[lastMeas].s.smoothed = S[lastMeas].s.filtered;
    S[lastMeas].s.smoothedCov = S[lastMeas].s.filteredCov;
    S[firstMeas].s.smoothed = R.res.states[rLast].s.filtered;
    S[firstMeas].s.smoothedCov = R.res.states[rLast].s.filteredCov;
    if (R.targetReached) {
      out.hasReference = true;
      out.referenceParams = R.targetParams;
      out.referenceCov = R.targetCov;
    }
  } else if (o.smoother == ACTS_KF_SMOOTHER_MBF) {
    mbfSmooth(in, S, tip);
  } else {
    // GainMatrixSmoother::operator() (GainMatrixSmoother.hpp:43-99)
    S[tip].s.smoothed = S[tip].s.filtered;
    S[tip].s.smoothedCov = S[tip].s.filteredCov;
    for (int k = tip - 1; k >= 0; --k) {
      int r = gainMatrixSmoothStep(S[k].s.filtered, S[k].s.filteredCov,
                                   S[k + 1].s.predicted, S[k + 1].s.predictedCov,
                                   S[k + 1].s.smoothed, S[k + 1].s.smoothedCov,
                                   S[k + 1].s.jacobian, S[k].s.smoothed,
                                   S[k].s.smoothedCov);
      if (r) {
        out.status = r;
        return;
      }
    }
  }

  // calculateTrackQuantities (TrackHelpers.hpp:421-445), tip backwards
  float chi2 = 0.f;
  for (int k = tip; k >= 0; --k) {
    const uint32_t f = S[k].s.flags;
    if (f & ACTS_KF_STATE_IS_HOLE) {
      ++out.nHoles;
    } else if (f & ACTS_KF_STATE_IS_OUTLIER) {
      ++out.nOutliers;
    } else if (f & ACTS_KF_STATE_HAS_MEASUREMENT) {
      ++out.nMeas;
      chi2 += S[k].s.chi2;
      out.ndf += in.meas(S[k].s.meas).dim;
    }
  }
  out.chi2 = chi2;
  out.states.reserve(tip + 1);
  for (int k = 0; k <= tip; ++k) out.states.push_back(S[k].s);
  out.status = 0;

  // KalmanFitter.hpp:959-971: with a reference surface the fit ends with
  // the extrapolation of the smoothed track, and fails if that fails (not
  // when the reverse filter reached the reference surface itself)
  if (o.reference_surface == ACTS_KF_REFERENCE_PERIGEE && !out.hasReference) {
    const int r = extrapolateToReference(geo, out);
    if (r) {
      const int nSteps = out.nSteps, nTrials = out.nTrials;
      out = TrackOut();
      out.status = r;
      out.nSteps = nSteps;
      out.nTrials = nTrials;
    }
  }
}

// findTrackStateForExtrapolation (TrackHelpers.hpp:180-320) and
// extrapolateTrackToReferenceSurface (TrackHelpers.hpp:339-383)
int extrapolateToReference(const Geometry& geo, TrackOut& out) {
  const std::vector<StateOut>& S = out.states;
  // findFirstMeasurementState / findLastMeasurementState: states whose type
  // is a measurement (not an outlier)
  int first = -1, last = -1;
  for (int k = 0; k < static_cast<int>(S.size()); ++k) {
    const uint32_t f = S[k].flags;
    if ((f & ACTS_KF_STATE_HAS_MEASUREMENT) && !(f & ACTS_KF_STATE_IS_OUTLIER)) {
      if (first < 0) first = k;
      last = k;
    }
  }
  if (first < 0) return ACTS_KF_STATUS_NO_COMPATIBLE_STATE;
  // the straight-line intersection of the reference surface from a state's
  // (smoothed) parameters
  auto intersect = [&](int k) {
    const Surface& s = geo.surfaces[S[k].surface];
    const Vec8 f = boundToFree(s, S[k].smoothed);
    const V3 pos = v3(f.a[0][0], f.a[1][0], f.a[2][0]);
    const V3 dir = v3(f.a[4][0], f.a[5][0], f.a[6][0]);
    return geo.reference.intersect(pos, dir, kOnSurfaceTolerance).is[0];
  };
  int k = -1;
  double distance = 0;
  const int strategy = geo.opt.reference_strategy;
  if (strategy == ACTS_KF_EXTRAPOLATE_FIRST || strategy == ACTS_KF_EXTRAPOLATE_LAST) {
    k = strategy == ACTS_KF_EXTRAPOLATE_FIRST ? first : last;
    const Intersection x = intersect(k);
    if (!x.valid()) return ACTS_KF_STATUS_REFERENCE_UNREACHABLE;
    distance = x.path;
  } else {
    const Intersection xf = intersect(first);
    const Intersection xl = intersect(last);
    const double absFirst = std::abs(xf.path), absLast = std::abs(xl.path);
    if (xf.valid() && absFirst <= absLast) {
      k = first;
      distance = xf.path;
    } else if (xl.valid() && absLast <= absFirst) {
      k = last;
      distance = xl.path;
    } else {
      return ACTS_KF_STATUS_REFERENCE_UNREACHABLE;
    }
  }
  // Direction::fromScalarZeroAsPositive
  const double direction = distance >= 0 ? 1.0 : -1.0;
  Vec6 endParams;
  M6 endCov, jacobian;
  double path = 0;
  int nSteps = 0;
  const int r = propagateToSurface(geo, S[k].smoothed, &S[k].smoothedCov,
                                   geo.surfaces[S[k].surface], geo.reference, direction,
                                   endParams, endCov, jacobian, path, nSteps,
                                   &geo.extrapolationOpt);
  if (r) return r;
  out.hasReference = true;
  out.referenceParams = endParams;
  out.referenceCov = endCov;
  return 0;
}

// ===========================================================================
// Propagator::propagate(start, target) with ForcedSurfaceReached
// (Propagator.ipp:237-257, StandardAborters.hpp:66-164) -- used by the
// oracle's own tests (helix closure, finite-difference jacobians).
// ===========================================================================
int propagateToSurface(const Geometry& geo, const Vec6& startParams,
                       const M6* startCov, const Surface& startSurface,
                       const Surface& target, double direction, Vec6& endParams,
                       M6& endCov, M6& jacobian, double& pathLength,
                       int& nSteps, const acts_kf_options* options) {
  const auto& o = options ? *options : geo.opt;
  Stepper st;
  initializeStepper(st, startSurface, startParams, startCov, o);
  double pathLimit = o.path_limit;
  setupLoopProtection(geo, o, st, pathLimit, false, direction);
  const double nearLimit = std::numeric_limits<double>::lowest();

  // SurfaceReached::checkAbort
  auto targetReached = [&]() -> bool {
    V3 d = st.direction();
    V3 dd = v3(direction * d[0], direction * d[1], direction * d[2]);
    MultiIntersection mi = target.intersect(st.position(), dd, o.surface_tolerance);
    // closest(): valid first, then smallest |path|
    int ci = 0;
    if (mi.size == 2) {
      const Intersection& a = mi.is[0];
      const Intersection& b = mi.is[1];
      bool bfirst;
      if (!a.valid() && b.valid())
        bfirst = true;
      else if (a.valid() && !b.valid())
        bfirst = false;
      else
        bfirst = std::abs(b.path) < std::abs(a.path);
      ci = bfirst ? 1 : 0;
    }
    bool reached = mi.is[ci].status == IStatus::onSurface;
    for (int i = 0; i < mi.size; ++i) {
      const Intersection& x = mi.is[i];
      if (x.valid() && x.path > nearLimit &&
          x.path < std::numeric_limits<double>::max() + kOnSurfaceTolerance) {
        // updateSurfaceStatus with the Actor constraint
        const Intersection& y = mi.is[i];
        if (y.status == IStatus::onSurface) {
          st.stepSize.release(ConstrainedStep::Actor);
          st.updateStepSize(y.path, ConstrainedStep::Actor);
        } else if (y.valid()) {
          st.stepSize.release(ConstrainedStep::Actor);
          st.updateStepSize(y.path, ConstrainedStep::Actor);
        }
        break;
      }
    }
    return reached;
  };
  auto checkAbort = [&]() -> bool {
    if (targetReached()) return true;
    return pathLimitCheckAbort(pathLimit, st, o.surface_tolerance);
  };

  bool terminated = checkAbort();
  unsigned steps = 0;
  int status = 0;
  for (; !terminated && steps < o.max_steps; ++steps) {
    StepResult sr = step(st, geo.field, geo.particle, o, direction);
    if (sr.status) {
      status = sr.status;
      break;
    }
    st.stepSize.release(ConstrainedStep::Navigator);
    st.stepSize.release(ConstrainedStep::Actor);
    if (checkAbort()) terminated = true;
  }
  if (!status && !terminated) status = ACTS_KF_STATUS_STEP_COUNT_LIMIT;
  nSteps = st.nSteps;
  if (status) return status;
  int r = boundState(target, st, startCov != nullptr, endParams);
  if (r) return r;
  endCov = st.cov;
  jacobian = st.jacobian;
  pathLength = st.pathAccumulated;
  return 0;
}

}  // namespace acts_oracle
