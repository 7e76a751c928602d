// kf_common.hpp -- geometry, field and material physics shared by the HIP
// fit kernels and the host-side event generator.
//
// Operation order follows the reference expressions term by term (Eigen
// evaluates the coefficient (i,j) of a fixed-size product as a k-ascending
// sum of products); structural zeros of the reference's dense matrices are
// skipped, which does not change any rounded result.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "../../include/acts_kf.h"

#define KF_HD __host__ __device__ inline

namespace kf {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2 * 3.14159265358979323846;
constexpr double kOnSurfaceTolerance = 1e-4;  // Definitions/Tolerance.hpp:25
constexpr double kNotSet = DBL_MAX;           // ConstrainedStep.hpp:163

// Physical constants as the reference evaluates its unit literals
// (Units.hpp: literal * unit in long double, rounded to double; the
// Interactions.cpp constants are then narrowed to float).  Computed on the
// host in long double and handed to the device in PhysConst.
struct PhysConst {
  float me;            // 0.5109989461 MeV   (Interactions.cpp:25)
  float k;             // 30.7075 MeV mm^2   (Interactions.cpp:27)
  float plasmaScale;   // 28.816 eV          (Interactions.cpp:29)
  float pad;
  double highland;     // 13.6 MeV           (Interactions.cpp:491)
  double rossiGreisen; // 17.5 MeV           (Interactions.cpp:500)
  double minP;         // 10 MeV             (PointwiseMaterialInteraction.hpp:199)
};

// Material slab of one surface after Material::fromMolarDensity and
// MaterialSlab construction (all float, Material.cpp, MaterialSlab.hpp).
struct DevSlab {
  float thickness, tInX0, tInL0;
  float molarElectronRho, meanI;
  int32_t vacuum;  // Material::isVacuum() (ar <= 0)
};

struct DevSurface {
  double R[9];    // local->global rotation, row-major
  double t[3];
  double Ri[9];   // Eigen general inverse of R (Transform::inverse, Affine)
  double ti[3];
  double b0, b1;  // bounds
  int32_t type;
  int32_t flags;
  int32_t mat;    // index into the slab table or -1
  int32_t pad;
};

// Packed symmetric 6x6 matrices: the upper triangle row by row (21 values).
KF_HD constexpr int sym(int i, int j) {
  return i <= j ? i * 6 - i * (i - 1) / 2 + (j - i) : j * 6 - j * (j - 1) / 2 + (i - j);
}

// ---------------------------------------------------------------------------
// vector helpers (Eigen coefficient order)
// ---------------------------------------------------------------------------
struct V3 {
  double x, y, z;
};
KF_HD V3 mk(double a, double b, double c) { return V3{a, b, c}; }
KF_HD double dot(const V3& a, const V3& b) {
  return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
KF_HD V3 cross(const V3& a, const V3& b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
KF_HD double perp(const V3& a) { return sqrt(a.x * a.x + a.y * a.y); }
KF_HD V3 normalized(const V3& a) {
  double z = dot(a, a);
  if (z > 0.0) {
    double n = sqrt(z);
    return mk(a.x / n, a.y / n, a.z / n);
  }
  return a;
}
// fastCathetus(hyp, arg) (MathHelpers.hpp:160-165)
KF_HD double cathetus(double hyp, double arg) {
  double h = fabs(arg);
  return sqrt((hyp - h) * (hyp + h));
}

// Eigen's 3x3 inverse via cofactors (compute_inverse_size3), used for the
// global->local transform of every surface (Transform3::inverse()).
KF_HD void invert3(const double* R, double* Ri) {
  auto m = [&](int i, int j) { return R[3 * i + j]; };
  auto cof = [&](int i, int j) {
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return m(i1, j1) * m(i2, j2) - m(i1, j2) * m(i2, j1);
  };
  double c0 = cof(0, 0), c1 = cof(1, 0), c2 = cof(2, 0);
  double det = (c0 * m(0, 0) + c1 * m(1, 0)) + c2 * m(2, 0);
  double invdet = 1.0 / det;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) Ri[3 * r + c] = cof(c, r) * invdet;
  Ri[0] = c0 * invdet;
  Ri[1] = c1 * invdet;
  Ri[2] = c2 * invdet;
}

KF_HD V3 toLocal(const DevSurface& s, const V3& p) {
  const double* Ri = s.Ri;
  return mk(((Ri[0] * p.x + Ri[1] * p.y) + Ri[2] * p.z) + s.ti[0],
            ((Ri[3] * p.x + Ri[4] * p.y) + Ri[5] * p.z) + s.ti[1],
            ((Ri[6] * p.x + Ri[7] * p.y) + Ri[8] * p.z) + s.ti[2]);
}
KF_HD V3 toGlobal(const double* R, const double* t, const V3& l) {
  return mk(((R[0] * l.x + R[1] * l.y) + R[2] * l.z) + t[0],
            ((R[3] * l.x + R[4] * l.y) + R[5] * l.z) + t[1],
            ((R[6] * l.x + R[7] * l.y) + R[8] * l.z) + t[2]);
}
// R * (l0, l1, 0) + t : the z term R[.,2]*0 only adds an exact zero
KF_HD V3 toGlobal2(const double* R, const double* t, double l0, double l1) {
  return mk((R[0] * l0 + R[1] * l1) + t[0], (R[3] * l0 + R[4] * l1) + t[1],
            (R[6] * l0 + R[7] * l1) + t[2]);
}
KF_HD V3 rotate(const double* R, const V3& v) {
  return mk((R[0] * v.x + R[1] * v.y) + R[2] * v.z,
            (R[3] * v.x + R[4] * v.y) + R[5] * v.z,
            (R[6] * v.x + R[7] * v.y) + R[8] * v.z);
}
KF_HD V3 col(const double* R, int j) { return mk(R[j], R[3 + j], R[6 + j]); }

// ---------------------------------------------------------------------------
// angles: periodic.hpp:19-93
// ---------------------------------------------------------------------------
KF_HD double wrapPeriodic(double value, double start, double range) {
  double diff = value - start;
  return ((0 <= diff) && (diff < range)) ? value
                                         : (value - range * floor(diff / range));
}
KF_HD double radianSym(double x) { return wrapPeriodic(x, -kPi, kTwoPi); }
KF_HD void normalizePhiTheta(double& phi, double& theta) {
  theta = wrapPeriodic(theta, 0.0, kTwoPi);
  if (kPi < theta) {
    phi += kPi;
    theta = kTwoPi - theta;
  }
  phi = radianSym(phi);
}

// ---------------------------------------------------------------------------
// Interactions.cpp (float arithmetic, RPP2018 eqs. 33.5/33.15)
// ---------------------------------------------------------------------------
struct RelQ {
  float q2OverBeta2, beta2, betaGamma, gamma;
};
KF_HD RelQ relQ(float mass, float qOverP, float absQ) {
  RelQ r;
  r.q2OverBeta2 = absQ * absQ + (mass * qOverP) * (mass * qOverP);
  const float mOverP = mass * fabsf(qOverP / absQ);
  const float pOverM = 1.0f / mOverP;
  r.beta2 = 1.0f / (1.0f + mOverP * mOverP);
  r.betaGamma = pOverM;
  r.gamma = sqrtf(1.0f * 1.0f + pOverM * pOverM);
  return r;
}
KF_HD float epsilonE(const PhysConst& c, float ne, float thickness, const RelQ& rq) {
  return 0.5f * c.k * ne * thickness * rq.q2OverBeta2;
}

// Natural log of a float, correctly rounded (through the double log), the
// canonical value of the reference's std::log(float).
KF_HD float flog(float x) { return static_cast<float>(log(static_cast<double>(x))); }

KF_HD float energyLossBethe(const PhysConst& c, const DevSlab& s, float m,
                            float qOverP, float absQ) {
  const RelQ rq = relQ(m, qOverP, absQ);
  const float I = s.meanI;
  const float ne = s.molarElectronRho;
  const float eps = epsilonE(c, ne, s.thickness, rq);
  float dhalf = 0.0f;
  if (!(rq.betaGamma < 10.0f)) {
    const float plasma = c.plasmaScale * sqrtf(ne / static_cast<float>(1 / 1000.0));
    dhalf = flog(rq.betaGamma) + flog(plasma / I) - 0.5f;
  }
  const float u = 2 * c.me * rq.betaGamma * rq.betaGamma;
  const float mfrac = c.me / m;
  const float nom = 2 * c.me * rq.betaGamma * rq.betaGamma;
  const float den = 1.0f + 2 * rq.gamma * mfrac + mfrac * mfrac;
  const float wmax = nom / den;
  const float running =
      flog(u / I) + flog(wmax / I) - 2.0f * rq.beta2 - 2.0f * dhalf;
  return eps * running;
}

KF_HD float energyLossLandauSigmaQOverP(const PhysConst& c, const DevSlab& s,
                                        float m, float qOverP, float absQ) {
  const RelQ rq = relQ(m, qOverP, absQ);
  const float fwhm = 4 * epsilonE(c, s.molarElectronRho, s.thickness, rq);
  const float sigmaE = fwhm * 0.42466090014400953f;
  const float pInv = qOverP / absQ;
  const float qOverBeta = sqrtf(rq.q2OverBeta2);
  return qOverBeta * pInv * pInv * sigmaE;
}

KF_HD float theta0(const PhysConst& c, const DevSlab& s, int absPdg, float m,
                   float qOverP, float absQ) {
  const float xOverX0 = s.tInX0;
  const float momentumInv = fabsf(qOverP / absQ);
  const float q2OverBeta2 = relQ(m, qOverP, absQ).q2OverBeta2;
  const float t = sqrtf(xOverX0 * q2OverBeta2);
  if (absPdg == 11)
    return static_cast<float>(c.rossiGreisen * momentumInv * t *
                              (1.0f + 0.125f * static_cast<float>(log10(static_cast<double>(10.0f * xOverX0)))));
  return static_cast<float>(c.highland * momentumInv * t *
                            (1.0f + 0.038f * 2 * flog(t)));
}

// ---------------------------------------------------------------------------
// field
// ---------------------------------------------------------------------------
// Magnetic field: constant, or an xyz grid as built by fieldMapXYZ
// (BFieldMapUtils.cpp:109-186): per axis nb bins of width `width` starting at
// `min`, bin b (1-based) holding the value of grid point b - 1; lookups are
// valid on [min, upper) with upper = the last grid point.
struct DevField {
  int32_t kind;
  int32_t pad;
  double B[3];
  int32_t nb[3];
  int32_t pad2;
  double min[3], width[3], upper[3];
  // The divisions of a lookup done as multiplications by reciprocals (set by
  // the host when the axes allow it, see buildGridField): rw = RN(1/width);
  // den = the upper minus the lower edge of a bin, as a lookup computes them,
  // the same for every bin of the axis, and rden = RN(1/den).  recip = 0:
  // divide.
  double rw[3], den[3], rden[3];
  int32_t recip;
  // values of points 0..nb per axis, the extra point nb (the overflow bin)
  // zero: [(ix * (nb1 + 1) + iy) * (nb2 + 1) + iz][3]
  const double* __restrict__ v;
  // the same map in float for the mixed-precision fit
  float minF[3], widthF[3], upperF[3];
  int32_t pad3;
  const float* __restrict__ vf;
  // Cell-major copies of the map on the device (null: use the point arrays
  // above): cell (ix, iy, iz), ix < nb0, iy < nb1, iz < nb2, at
  // (ix * nb1 + iy) * nb2 + iz holds the values of its 8 corner points in
  // the order the interpolation reads them: corner (dx, dy) = (i >> 1, i & 1)
  // for i = 0..3, then dz = 0, 1, then the 3 components.  A lookup reads one
  // contiguous record (kCellDoubles doubles, kCellFloats floats with 8 of
  // padding) instead of four point pairs in different cache lines.
  const double* __restrict__ cells;
  const float* __restrict__ cellsF;
  // The lane's copy of the last cell record it read (the reference's cached
  // FieldCell, InterpolatedBFieldMap.hpp:354-366), in LDS, value j at
  // cache[j * kCellCacheStride] (cacheF for the FP32 map), and that cell's
  // index at cacheIdx[0]; null (host code, kernels without it): no cache.
  double* cache;
  float* cacheF;
  int64_t* cacheIdx;
};
// a / b rounded to nearest, from y = RN(1/b): q = a * y, then two
// corrections q + (a - b q) y with fused multiply-adds.  The first makes q a
// faithful quotient; from a faithful q and a reciprocal within half an ulp
// the second gives the correctly rounded quotient (Markstein's theorem,
// IBM J. Res. Dev. 34 (1990) 111; Muller et al., Handbook of Floating-Point
// Arithmetic, 2018, sec. 4.9), i.e. exactly a / b, for normal b, y and a
// quotient far from overflow and underflow.  Five full-rate operations in
// place of the division's scale, reciprocal, refinement and fix-up sequence
// (2 % off the FP64 field-map forward kernel; the same for the FP32 map's
// divisions measured 4 % slower than the division, so it divides).
KF_HD double divRecip(double a, double b, double y) {
  double q = a * y;
  q = fma(fma(-q, b, a), y, q);
  return fma(fma(-q, b, a), y, q);
}

constexpr int kCellValues = 24;
constexpr int kCellDoubles = 24;  // 192 B: two 128-byte cache lines
constexpr int kCellFloats = 32;   // 128 B: one cache line
constexpr int kCellCacheStride = 64;  // lanes of the wavefront sharing the LDS cache

}  // namespace kf
