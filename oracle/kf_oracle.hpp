// kf_oracle.hpp -- CPU restatement of the reference Kalman-fit path.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP fitter is checked
// against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it, and only as the checker / the timed CPU baseline.  The
// product (acts_amd/) never links or calls it.
//
// It restates, in plain C++ with no Eigen, the reference functions on the
// path KalmanFitter::fit -> Propagator<EigenStepper, DirectNavigator>:
// every dense matrix expression is evaluated in the order Eigen 3.4 evaluates
// it for fixed-size column-major matrices on SSE2 (coefficient (i,j) of A*B
// accumulated as sum_k A(i,k)*B(k,j) with k ascending, no FMA), and the
// reference's float islands (material physics, q/p narrowing, chi2 storage)
// are kept as float.  Each function cites the reference file:line it follows.
//
// Parity pinning (see tests/test_oracle_*.py): GainMatrixUpdater and
// GainMatrixSmoother against the reference unit-test known answers;
// Bethe/Landau/Highland against the reference's PDG checks and the RPP
// closed forms; the stepper against the analytic helix; the transport
// jacobian against finite differences of the restated propagator; surfaces
// against closed-form intersections and local/global round trips; the fitter
// against the reference KalmanFitter test invariants and pull widths.
// Last-bit agreement with Eigen's PartialPivLU inverse and with glibc's
// transcendental functions as linked into the reference is parity unpinned.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "../include/acts_kf.h"

namespace acts_oracle {

constexpr double kPi = 3.14159265358979323846;
constexpr double kOnSurfaceTolerance = 1e-4;  // Definitions/Tolerance.hpp:25
constexpr double kNotSet = std::numeric_limits<double>::max();

// ---------------------------------------------------------------------------
// small dense matrices (row-major storage, Eigen evaluation order)
// ---------------------------------------------------------------------------
template <int R, int C>
struct Mat {
  double a[R][C];
  double& operator()(int i, int j) { return a[i][j]; }
  double operator()(int i, int j) const { return a[i][j]; }
  static Mat zero() {
    Mat m;
    for (int i = 0; i < R; ++i)
      for (int j = 0; j < C; ++j) m.a[i][j] = 0.0;
    return m;
  }
  static Mat identity() {
    Mat m = zero();
    for (int i = 0; i < (R < C ? R : C); ++i) m.a[i][i] = 1.0;
    return m;
  }
};

template <int R, int K, int C>
inline Mat<R, C> mul(const Mat<R, K>& A, const Mat<K, C>& B) {
  Mat<R, C> out;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) {
      double acc = A.a[i][0] * B.a[0][j];
      for (int k = 1; k < K; ++k) acc = acc + A.a[i][k] * B.a[k][j];
      out.a[i][j] = acc;
    }
  return out;
}

// A * B^T
template <int R, int K, int C>
inline Mat<R, C> mulT(const Mat<R, K>& A, const Mat<C, K>& B) {
  Mat<R, C> out;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) {
      double acc = A.a[i][0] * B.a[j][0];
      for (int k = 1; k < K; ++k) acc = acc + A.a[i][k] * B.a[j][k];
      out.a[i][j] = acc;
    }
  return out;
}

template <int R, int C>
inline Mat<C, R> transpose(const Mat<R, C>& A) {
  Mat<C, R> out;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) out.a[j][i] = A.a[i][j];
  return out;
}

template <int R, int C>
inline bool hasNaN(const Mat<R, C>& A) {
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j)
      if (std::isnan(A.a[i][j])) return true;
  return false;
}

struct V3 {
  double x[3];
  double operator[](int i) const { return x[i]; }
  double& operator[](int i) { return x[i]; }
};

inline V3 v3(double a, double b, double c) { return V3{{a, b, c}}; }
// Eigen cross product (Geometry/OrthoMethods.h)
inline V3 cross(const V3& a, const V3& b) {
  return v3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
            a[0] * b[1] - a[1] * b[0]);
}
inline double dot(const V3& a, const V3& b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
inline double squaredNorm(const V3& a) { return dot(a, a); }
inline double norm(const V3& a) { return std::sqrt(squaredNorm(a)); }
inline V3 normalized(const V3& a) {
  double z = squaredNorm(a);
  if (z > 0.0) {
    double n = std::sqrt(z);
    return v3(a[0] / n, a[1] / n, a[2] / n);
  }
  return a;
}
inline double perp(const V3& a) { return std::sqrt(a[0] * a[0] + a[1] * a[1]); }
inline double phiOf(const V3& a) { return std::atan2(a[1], a[0]); }
inline double thetaOf(const V3& a) { return std::atan2(perp(a), a[2]); }

// MathHelpers.hpp:94-180: fastHypot is a right fold of squares.
inline double fastCathetus1(double hyp, double arg) {
  double h = std::abs(arg);
  return std::sqrt((hyp - h) * (hyp + h));
}

// ---------------------------------------------------------------------------
// 3x3 affine placement with Eigen's general inverse (Transform::inverse with
// the Affine hint -> compute_inverse_size3 on the linear part)
// ---------------------------------------------------------------------------
struct Placement {
  double R[3][3];   // local -> global rotation (columns = local axes)
  double t[3];
  double Ri[3][3];  // inverse linear part
  double ti[3];

  void set(const double* rot9, const double* t3) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) R[i][j] = rot9[3 * i + j];
    for (int i = 0; i < 3; ++i) t[i] = t3[i];
    auto cof = [&](int i, int j) {
      int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      return R[i1][j1] * R[i2][j2] - R[i1][j2] * R[i2][j1];
    };
    double c0[3] = {cof(0, 0), cof(1, 0), cof(2, 0)};
    double det = (c0[0] * R[0][0] + c0[1] * R[1][0]) + c0[2] * R[2][0];
    double invdet = 1.0 / det;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) Ri[r][c] = cof(c, r) * invdet;
    for (int j = 0; j < 3; ++j) Ri[0][j] = c0[j] * invdet;
    for (int i = 0; i < 3; ++i)
      ti[i] = -((Ri[i][0] * t[0] + Ri[i][1] * t[1]) + Ri[i][2] * t[2]);
  }
  V3 toGlobal(const V3& l) const {
    V3 g;
    for (int i = 0; i < 3; ++i)
      g[i] = ((R[i][0] * l[0] + R[i][1] * l[1]) + R[i][2] * l[2]) + t[i];
    return g;
  }
  V3 toLocal(const V3& p) const {
    V3 l;
    for (int i = 0; i < 3; ++i)
      l[i] = ((Ri[i][0] * p[0] + Ri[i][1] * p[1]) + Ri[i][2] * p[2]) + ti[i];
    return l;
  }
  V3 rotate(const V3& v) const {
    V3 g;
    for (int i = 0; i < 3; ++i)
      g[i] = (R[i][0] * v[0] + R[i][1] * v[1]) + R[i][2] * v[2];
    return g;
  }
  V3 col(int j) const { return v3(R[0][j], R[1][j], R[2][j]); }
};

// ---------------------------------------------------------------------------
// Material (Material.cpp, MaterialSlab.hpp) -- all float
// ---------------------------------------------------------------------------
struct MaterialF {
  float x0 = std::numeric_limits<float>::infinity();
  float l0 = std::numeric_limits<float>::infinity();
  float ar = 0.f, z = 0.f, molarRho = 0.f, molarElectronRho = 0.f, meanI = 0.f;
  bool isVacuum() const { return ar <= 0.f; }
};

MaterialF materialFromMolarDensity(float x0, float l0, float ar, float z,
                                   float molarRho);

struct SlabF {
  MaterialF mat;
  float thickness = 0.f, tInX0 = 0.f, tInL0 = 0.f;
  bool isVacuum() const { return mat.isVacuum() || thickness <= 0.f; }
  static SlabF make(const MaterialF& m, float thickness);
  void scaleThickness(float s) {
    thickness *= s;
    tInX0 *= s;
    tInL0 *= s;
  }
};

struct Particle {
  int absPdg = 13;
  float mass = 0.f;
  float absQ = 1.f;
  double extractMomentum(double qop) const {
    return static_cast<double>(static_cast<float>(std::copysign(absQ, qop))) / qop;
  }
  double qOverP(double p, float signedQ) const { return signedQ / p; }
};

// Interactions.cpp (float)
float energyLossBethe(const SlabF& slab, float m, float qOverP, float absQ);
float energyLossLandau(const SlabF& slab, float m, float qOverP, float absQ);
float energyLossLandauFwhm(const SlabF& slab, float m, float qOverP, float absQ);
float energyLossLandauSigmaQOverP(const SlabF& slab, float m, float qOverP,
                                  float absQ);
float multipleScatteringTheta0(const SlabF& slab, int absPdg, float m,
                               float qOverP, float absQ);

// ---------------------------------------------------------------------------
// Surfaces (Surface.cpp, CylinderSurface.cpp, DiscSurface.cpp,
// PlaneSurface.cpp, PlanarHelper.hpp, RealQuadraticEquation.hpp)
// ---------------------------------------------------------------------------
enum class IStatus { unreachable = 0, reachable = 1, onSurface = 2 };

struct Intersection {
  double path = std::numeric_limits<double>::infinity();
  IStatus status = IStatus::unreachable;
  bool valid() const { return status != IStatus::unreachable; }
};
struct MultiIntersection {
  Intersection is[2];
  int size = 1;
};

struct Surface {
  int type = ACTS_KF_SURFACE_PLANE;
  Placement pl;
  double b0 = 0, b1 = 0;       // bounds
  bool sensitive = false;
  bool hasMaterial = false;
  SlabF slab;                  // homogeneous material, split factor 1

  MultiIntersection intersect(const V3& pos, const V3& dir, double tol) const;
  bool globalToLocal(const V3& pos, double& l0, double& l1,
                     double tol = kOnSurfaceTolerance) const;
  V3 localToGlobal(double l0, double l1) const;
  Mat<3, 3> referenceFrame(const V3& pos) const;
  // the same with the direction, which a LineSurface's frame depends on
  // (LineSurface.cpp:57-114); the other surface types ignore it
  bool globalToLocal(const V3& pos, const V3& dir, double& l0, double& l1,
                     double tol = kOnSurfaceTolerance) const;
  V3 localToGlobal(double l0, double l1, const V3& dir) const;
  Mat<3, 3> referenceFrame(const V3& pos, const V3& dir) const;
  double pathCorrection(const V3& pos, const V3& dir) const;
  Mat<8, 6> boundToFreeJacobian(const V3& pos, const V3& dir) const;
  Mat<6, 8> freeToBoundJacobian(const V3& pos, const V3& dir) const;
  Mat<1, 8> freeToPathDerivative(const V3& pos, const V3& dir) const;
  V3 cylinderNormal(const V3& pos) const;
};

using Vec6 = Mat<6, 1>;
using Vec8 = Mat<8, 1>;
using M6 = Mat<6, 6>;
using M8 = Mat<8, 8>;

// TransformationHelpers.cpp:17-63
Vec8 boundToFree(const Surface& s, const Vec6& b);
int freeToBound(const Vec8& f, const Surface& s, Vec6& out);

// periodic.hpp / TrackParameterHelpers.hpp
double wrapPeriodic(double value, double start, double range);
double radianPos(double x);
double radianSym(double x);
void normalizePhiTheta(double& phi, double& theta);

// ---------------------------------------------------------------------------
// Magnetic field
// ---------------------------------------------------------------------------
struct FieldGrid {
  // InterpolatedBFieldMap over an xyz Grid built by fieldMapXYZ
  // (BFieldMapUtils.cpp:109-186, Helpers.hpp:289-310): per axis, `nBins`
  // equidistant bins on [min, max) with the upper edge one step past the last
  // point, plus under/overflow bins; values are defined on bin lower edges.
  int nb[3];        // number of bins per axis (= number of points)
  double min[3], max[3], width[3];
  std::vector<double> v;  // values at grid points [(i*nb1 + j)*nb2 + k][3]
  void set(const int n[3], const double pmin[3], const double pmax[3], const double* values);
  int get(const V3& pos, V3& out) const;
};

void ellipticKE(double k, double& K, double& E);

// SolenoidBField with its Config (radius, length, nCoils, bMagCenter)
struct Solenoid {
  double radius, length;
  int nCoils;
  double dz, R2, scale;
  Solenoid(double radius, double length, int nCoils, double bMagCenter);
  double k2(double r, double z) const;
  double Br(double r, double z, double scale) const;
  double Bz(double r, double z, double scale) const;
  void multiCoil(double r, double z, double scale, double& br, double& bz) const;
  V3 field(const V3& pos) const;
};

struct Field {
  int kind = ACTS_KF_FIELD_CONSTANT;
  double B[3] = {0, 0, 0};
  FieldGrid grid;
  // returns status (0 ok, FIELD_OUT_OF_BOUNDS)
  int get(const V3& pos, V3& out) const;
};

// ---------------------------------------------------------------------------
// Fitter options and outputs
// ---------------------------------------------------------------------------
struct Geometry {
  std::vector<Surface> surfaces;
  Field field;
  acts_kf_options opt;
  Particle particle;
  // the default propagator and stepper options, which the extrapolation to
  // the reference surface uses (KalmanFitter.hpp:960-961)
  acts_kf_options extrapolationOpt;
  Surface reference;  // the reference surface (ACTS_KF_SURFACE_LINE)
};

struct StateOut {
  int surface = -1;
  int meas = -1;                    // index of the state's measurement
  uint32_t flags = 0;
  double path = 0;
  float chi2 = 0.f;
  Vec6 predicted, filtered, smoothed;
  M6 predictedCov, filteredCov, smoothedCov, jacobian;
  bool isMeasurementState = false;  // created in the measurement branch
};

struct TrackOut {
  int status = 0;
  bool hasReference = false;
  bool reversed = false;  // smoothed by reverse filtering
  Vec6 referenceParams;
  M6 referenceCov;
  std::vector<StateOut> states;  // up to and including the tip
  float chi2 = 0.f;
  int ndf = 0, nMeas = 0, nHoles = 0, nOutliers = 0;
  int nSteps = 0, nTrials = 0;
};

// One measurement: N = dim in {1, 2} bound parameters idx[0..N-1] with
// the values v[0..N-1] and the N x N row-major covariance V.
struct Meas {
  int dim = 2;
  int idx[2] = {0, 1};
  double v[2] = {0, 0};
  double V[4] = {0, 0, 0, 0};
};

struct TrackIn {
  Vec6 startParams;
  M6 startCov;
  Surface startSurface;
  std::vector<int> sequence;
  std::vector<int> measSurface;
  std::vector<double> measValue;  // 2 per measurement
  std::vector<double> measCov;    // 4 per measurement
  std::vector<uint8_t> measDim;   // per measurement (empty: all 2)
  std::vector<uint8_t> measIdx;   // 2 per measurement (empty: 0, 1)
  Meas meas(int i) const {
    Meas m;
    m.dim = measDim.empty() ? 2 : measDim[i];
    if (!measIdx.empty()) {
      m.idx[0] = measIdx[2 * i];
      m.idx[1] = measIdx[2 * i + 1];
    }
    for (int k = 0; k < 2; ++k) m.v[k] = measValue[2 * i + k];
    for (int k = 0; k < 4; ++k) m.V[k] = measCov[4 * i + k];
    return m;
  }
};

// KalmanFitter::fit with DirectNavigator + GainMatrixUpdater +
// GainMatrixSmoother (KalmanFitter.hpp:721-984)
void fitTrack(const Geometry& geo, const TrackIn& in, TrackOut& out);

// GainMatrixUpdaterImpl.hpp:21-91 (N = 2, projector on loc0/loc1)
int gainMatrixUpdate(const Vec6& predicted, const M6& predictedCov,
                     const double meas[2], const double measCov[4],
                     bool useJoseph, Vec6& filtered, M6& filteredCov,
                     double& chi2);
// the same for a measurement of dimension 1 or 2 on any subspace
int gainMatrixUpdate(const Vec6& predicted, const M6& predictedCov, const Meas& m,
                     bool useJoseph, Vec6& filtered, M6& filteredCov, double& chi2);

// GainMatrixSmoother.cpp:18-97 for one pair of states
int gainMatrixSmoothStep(const Vec6& filtered, const M6& filteredCov,
                         const Vec6& prevPredicted, const M6& prevPredictedCov,
                         const Vec6& prevSmoothed, const M6& prevSmoothedCov,
                         const M6& prevJacobian, Vec6& smoothed,
                         M6& smoothedCov);

M6 inverse6(const M6& A);

// Propagator::propagate(start, target, options) with covariance transport
// and the ForcedSurfaceReached aborter (StandardAborters.hpp:161-164);
// returns status and the bound parameters / covariance / jacobian on target.
int propagateToSurface(const Geometry& geo, const Vec6& startParams,
                       const M6* startCov, const Surface& startSurface,
                       const Surface& target, double direction, Vec6& endParams,
                       M6& endCov, M6& jacobian, double& pathLength,
                       int& nSteps, const acts_kf_options* options = nullptr);

// extrapolateTrackToReferenceSurface (TrackHelpers.hpp:180-383) for a fitted
// track: the state chosen by the strategy, its smoothed parameters, to the
// reference surface; returns a status
int extrapolateToReference(const Geometry& geo, TrackOut& out);

// The per-state columns of RootTrackStatesWriter for a fitted batch, in
// the layout of acts_kf_track_states_out (track_states.cpp).
void trackStatesTable(const Geometry& geo, const acts_kf_batch_in& in,
                      const acts_kf_batch_out& fitted, const double* truthMeas,
                      const acts_kf_track_states_out& out);

}  // namespace acts_oracle
