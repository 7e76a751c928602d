// kf_mixed.hpp -- the FP32 stepping of the mixed-precision fit.
//
// Same algorithm as rkStep (EigenStepper::step, EigenStepper.ipp:169-388,
// and the transport matrix of EigenStepperDefaultExtension.hpp:137-228), with
// the Runge-Kutta stages, the error estimate, the field lookups (FP32 copy
// of the field map, FP32 trilinear interpolation) and the transport-matrix
// update computed in float.  The track state (position, direction, q/p,
// time, path length, jacobian blocks) stays in double: each step's float
// increments are added to it, so rounding does not accumulate at float
// precision over a track.  The step length, the straight-line part h*dir of
// the position update and the direction's renormalisation are double too:
// they set where a track lands on a surface, and their float rounding would
// show up as ~1e-5 relative q/p error at 100 GeV.  So are the time, dt/ds
// and d(t)/d(q/p).  Everything per surface (covariance transport, gain-matrix
// update, material, smoother) stays FP64.
#pragma once

#include "kf_device.hpp"

namespace kf {

struct F3 {
  float x, y, z;
};
__device__ inline F3 mkf(float a, float b, float c) { return F3{a, b, c}; }
__device__ inline F3 tof(const V3& v) {
  return F3{static_cast<float>(v.x), static_cast<float>(v.y), static_cast<float>(v.z)};
}
__device__ inline F3 crossf(const F3& a, const F3& b) {
  return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Trilinear lookup in the FP32 map (same cell, corner order and collapse
// order as gridField).
__device__ inline int gridFieldF(const DevField& F, const F3& p, F3& out) {
  const float x[3] = {p.x, p.y, p.z};
  int lo[3];
  float f[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float mn = F.minF[a], w = F.widthF[a];
    if (!(x[a] >= mn) || x[a] >= F.upperF[a]) return ACTS_KF_STATUS_FIELD_OUT_OF_BOUNDS;
    int bin = static_cast<int>(floorf((x[a] - mn) / w) + 1.f);
    bin = bin < 1 ? 1 : (bin > F.nb[a] ? F.nb[a] : bin);
    lo[a] = bin - 1;
    const float ll = mn + static_cast<float>(bin - 1) * w;
    f[a] = (x[a] - ll) / w;
  }
  float c[4][3];
  if (F.cellsF) {
    // the cell's record: one 128-byte line
    const int64_t cell = (static_cast<int64_t>(lo[0]) * F.nb[1] + lo[1]) * F.nb[2] + lo[2];
    float v[kCellValues];
    if (F.cacheF && F.cacheIdx[0] == cell) {
#pragma unroll
      for (int j = 0; j < kCellValues; ++j) v[j] = F.cacheF[j * kCellCacheStride];
    } else {
      const float* q = static_cast<const float*>(
          __builtin_assume_aligned(F.cellsF + kCellFloats * cell, 128));
#pragma unroll
      for (int j = 0; j < kCellValues; ++j) v[j] = q[j];
      if (F.cacheF) {
#pragma unroll
        for (int j = 0; j < kCellValues; ++j) F.cacheF[j * kCellCacheStride] = v[j];
        F.cacheIdx[0] = cell;
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        c[i][k] = (1.f - f[2]) * v[6 * i + k] + f[2] * v[6 * i + 3 + k];
  } else {
    const int64_t sy = F.nb[2] + 1;
    const int64_t sx = static_cast<int64_t>(F.nb[1] + 1) * (F.nb[2] + 1);
    const float* base = F.vf + 3 * (lo[0] * sx + lo[1] * sy + lo[2]);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float* q = base + 3 * (((i >> 1) & 1) * sx + (i & 1) * sy);
#pragma unroll
      for (int k = 0; k < 3; ++k) c[i][k] = (1.f - f[2]) * q[k] + f[2] * q[3 + k];
    }
  }
  float d[2][3];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) d[i][k] = (1.f - f[1]) * c[2 * i][k] + f[1] * c[2 * i + 1][k];
  out = mkf((1.f - f[0]) * d[0][0] + f[0] * d[1][0], (1.f - f[0]) * d[0][1] + f[0] * d[1][1],
            (1.f - f[0]) * d[0][2] + f[0] * d[1][2]);
  return 0;
}

__device__ inline int fieldAtF(const DevField& F, const F3& p, F3& out) {
  if (F.kind == ACTS_KF_FIELD_CONSTANT) {
    out = mkf(static_cast<float>(F.B[0]), static_cast<float>(F.B[1]),
              static_cast<float>(F.B[2]));
    return 0;
  }
  return gridFieldF(F, p, out);
}

__device__ inline float stepScalingF(float tol, float err) {
  float x = tol / err;
  x = sqrtf(sqrtf(x));
  return fminf(fmaxf(x, 0.25f), 4.0f);
}

__device__ inline int rkStepMixed(Track& T, const DevField& F, const DevOpts& o,
                                  double propDir = 1.0, bool transport = true) {
  const F3 pos = tof(T.pos);
  const F3 dir = tof(T.dir);
  const float qop = static_cast<float>(T.qop);
  const float tol = static_cast<float>(o.stepTolerance);
  F3 B1;
  if (const int fs = fieldAtF(F, pos, B1)) return fs;
  const F3 c1 = crossf(dir, B1);
  const F3 k1 = mkf(qop * c1.x, qop * c1.y, qop * c1.z);

  const double initialH = stepValue(T) * propDir;
  // the step length actually taken, kept in double: a step the navigator
  // sized to reach a surface must land on it to double precision
  double hD = initialH;
  float h = static_cast<float>(initialH);
  float h2, half_h, err;
  F3 B2, B3, k2, k3, k4;
  int nTrials = 0;
  while (true) {
    ++nTrials;
    h2 = h * h;
    half_h = h * 0.5f;
    const float h28 = h2 * 0.125f;
    const F3 p1 = mkf((pos.x + half_h * dir.x) + h28 * k1.x, (pos.y + half_h * dir.y) + h28 * k1.y,
                      (pos.z + half_h * dir.z) + h28 * k1.z);
    if (const int fs = fieldAtF(F, p1, B2)) return fs;
    F3 a = crossf(mkf(dir.x + half_h * k1.x, dir.y + half_h * k1.y, dir.z + half_h * k1.z), B2);
    k2 = mkf(qop * a.x, qop * a.y, qop * a.z);
    a = crossf(mkf(dir.x + half_h * k2.x, dir.y + half_h * k2.y, dir.z + half_h * k2.z), B2);
    k3 = mkf(qop * a.x, qop * a.y, qop * a.z);
    const float h22 = h2 * 0.5f;
    const F3 p2 = mkf((pos.x + h * dir.x) + h22 * k3.x, (pos.y + h * dir.y) + h22 * k3.y,
                      (pos.z + h * dir.z) + h22 * k3.z);
    if (const int fs = fieldAtF(F, p2, B3)) return fs;
    a = crossf(mkf(dir.x + h * k3.x, dir.y + h * k3.y, dir.z + h * k3.z), B3);
    k4 = mkf(qop * a.x, qop * a.y, qop * a.z);
    const float e0 = ((k1.x - k2.x) - k3.x) + k4.x;
    const float e1 = ((k1.y - k2.y) - k3.y) + k4.y;
    const float e2 = ((k1.z - k2.z) - k3.z) + k4.z;
    err = h2 * ((fabsf(e0) + fabsf(e1)) + fabsf(e2));
    err = fmaxf(1e-20f, err);
    if (err <= 4.0f * tol) break;
    h *= stepScalingF(tol, err);
    hD = static_cast<double>(h);
    if (fabs(hD) < fabs(o.stepSizeCutOff))
      return ACTS_KF_STATUS_STEP_SIZE_STALLED;
    if (static_cast<uint32_t>(nTrials) > o.maxTrials) return ACTS_KF_STATUS_STEP_ADJUSTMENT_FAILED;
  }

  // time and its derivatives in double (as rkTime): the time is not
  // measured, so its smoothed value is set by these jacobian entries through
  // the ill-conditioned RTS gain, which magnifies float rounding to ~1e-5
  const float m = o.mass;
  const double pD = copysign(static_cast<double>(o.absQ), T.qop) / T.qop;
  const double mpD = m / pD;
  const double dtds = sqrt(1 + mpD * mpD);
  T.time += hD * dtds;
  T.der[3] = dtds;

  const float h6 = h / 6.f;
  // transport matrix D (only the non-trivial blocks); skipped without
  // `transport` (the trajectory does not depend on it)
  if (transport) {
  const float hh6 = (h * h) / 6.f;
  const float qh2 = qop * half_h;
  const float qh = qop * h;
  const F3 dk1dL = crossf(dir, B1);
  F3 dk2dL, dk3dL, dk4dL;
  {
    const F3 a = crossf(mkf(dir.x + half_h * k1.x, dir.y + half_h * k1.y, dir.z + half_h * k1.z), B2);
    const F3 b = crossf(dk1dL, B2);
    dk2dL = mkf(a.x + qh2 * b.x, a.y + qh2 * b.y, a.z + qh2 * b.z);
  }
  {
    const F3 a = crossf(mkf(dir.x + half_h * k2.x, dir.y + half_h * k2.y, dir.z + half_h * k2.z), B2);
    const F3 b = crossf(dk2dL, B2);
    dk3dL = mkf(a.x + qh2 * b.x, a.y + qh2 * b.y, a.z + qh2 * b.z);
  }
  {
    const F3 a = crossf(mkf(dir.x + h * k3.x, dir.y + h * k3.y, dir.z + h * k3.z), B3);
    const F3 b = crossf(dk3dL, B3);
    dk4dL = mkf(a.x + qh * b.x, a.y + qh * b.y, a.z + qh * b.z);
  }
  float dk1dT[9], dk2dT[9], dk3dT[9], dk4dT[9];
  dk1dT[0] = 0.f;          dk1dT[1] = qop * B1.z;  dk1dT[2] = qop * -B1.y;
  dk1dT[3] = qop * -B1.z;  dk1dT[4] = 0.f;         dk1dT[5] = qop * B1.x;
  dk1dT[6] = qop * B1.y;   dk1dT[7] = qop * -B1.x; dk1dT[8] = 0.f;
  auto stage = [&](const float* prev, float hh, const F3& B, float* out) {
    float m3[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) m3[3 * i + j] = (i == j ? 1.f : 0.f) + hh * prev[3 * i + j];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const F3 c = crossf(mkf(m3[j], m3[3 + j], m3[6 + j]), B);
      out[j] = qop * c.x;
      out[3 + j] = qop * c.y;
      out[6 + j] = qop * c.z;
    }
  };
  stage(dk1dT, half_h, B2, dk2dT);
  stage(dk2dT, half_h, B2, dk3dT);
  stage(dk3dT, h, B3, dk4dT);

  float dF[12], dG[12];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int e = 3 * i + j;
      const float id = (i == j ? 1.f : 0.f);
      dF[4 * i + j] = (id + h6 * ((dk1dT[e] + dk2dT[e]) + dk3dT[e])) * h;
      dG[4 * i + j] = id + h6 * ((dk1dT[e] + 2.f * (dk2dT[e] + dk3dT[e])) + dk4dT[e]);
    }
  }
  dF[3] = hh6 * ((dk1dL.x + dk2dL.x) + dk3dL.x);
  dF[7] = hh6 * ((dk1dL.y + dk2dL.y) + dk3dL.y);
  dF[11] = hh6 * ((dk1dL.z + dk2dL.z) + dk3dL.z);
  dG[3] = h6 * ((dk1dL.x + 2.f * (dk2dL.x + dk3dL.x)) + dk4dL.x);
  dG[7] = h6 * ((dk1dL.y + 2.f * (dk2dL.y + dk3dL.y)) + dk4dL.y);
  dG[11] = h6 * ((dk1dL.z + 2.f * (dk2dL.z + dk3dL.z)) + dk4dL.z);
  const double D37 = hD * m * m / (pD * pD * T.qop * dtds);

  float br[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) br[i] = static_cast<float>(T.br[i]);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float pt = (dF[4 * i + 0] * br[j] + dF[4 * i + 1] * br[4 + j]) + dF[4 * i + 2] * br[8 + j];
      float pb = (dG[4 * i + 0] * br[j] + dG[4 * i + 1] * br[4 + j]) + dG[4 * i + 2] * br[8 + j];
      if (j == 3) {
        pt = pt + dF[4 * i + 3];
        pb = pb + dG[4 * i + 3];
      }
      T.tr[4 * i + j] += static_cast<double>(pt);
      T.br[4 * i + j] = static_cast<double>(pb);
    }
  }
  T.tr37 += D37;
  }

  // the straight-line part h*dir in double (in float it would carry an
  // absolute error of |h| * 6e-8, i.e. 6e-5 mm for a 1 m step, a 1e-5
  // relative error of q/p at 100 GeV); the bending part in float
  const float h26 = h2 / 6.f;
  T.pos.x += hD * T.dir.x + static_cast<double>(h26 * ((k1.x + k2.x) + k3.x));
  T.pos.y += hD * T.dir.y + static_cast<double>(h26 * ((k1.y + k2.y) + k3.y));
  T.pos.z += hD * T.dir.z + static_cast<double>(h26 * ((k1.z + k2.z) + k3.z));
  const V3 nd = mk(T.dir.x + static_cast<double>(h6 * ((k1.x + 2.f * (k2.x + k3.x)) + k4.x)),
                   T.dir.y + static_cast<double>(h6 * ((k1.y + 2.f * (k2.y + k3.y)) + k4.y)),
                   T.dir.z + static_cast<double>(h6 * ((k1.z + 2.f * (k2.z + k3.z)) + k4.z)));
  // renormalise: float reciprocal square root, one Newton step in double
  const double nn = dot(nd, nd);
  double inv = static_cast<double>(rsqrtf(static_cast<float>(nn)));
  inv = inv * (1.5 - (0.5 * nn) * (inv * inv));
  T.dir = mk(nd.x * inv, nd.y * inv, nd.z * inv);
  T.der[0] = T.dir.x;
  T.der[1] = T.dir.y;
  T.der[2] = T.dir.z;
  T.der[4] = k4.x;
  T.der[5] = k4.y;
  T.der[6] = k4.z;
  T.pathAcc += hD;
  ++T.nSteps;
  T.nTrials += nTrials;

  const double nextAccuracy = fabs(hD * static_cast<double>(stepScalingF(tol, err)));
  const double previousAccuracy = fabs(T.cAcc);
  const double initialStepLength = fabs(initialH);
  if (nextAccuracy < initialStepLength || nextAccuracy > previousAccuracy) T.cAcc = nextAccuracy;
  return 0;
}

// The SympyStepper's step (rkStepSympy) in mixed precision: the stages,
// the error estimate and the jacobian recursion in float; the straight-line
// parts h dir of position and direction, the time, dt/ds, d(time)/d(q/p)
// and the renormalisation in double, as in rkStepMixed.  In float the
// error estimate is summed from the small stage terms,
// kick1 - c2 - c3 + kick4 = h/2 (k1 - k2 - k3 + k4) with c_i the bends of
// dir2 and dir3, and the q/p column's field terms dir3 - dir and dir4 - dir
// are c2 and c3: the FP64 form's differences of O(1) vectors would leave
// float noise of ~1e-7 in them.
__device__ inline int rkStepSympyMixed(Track& T, const DevField& F, const DevOpts& o,
                                       double propDir = 1.0, bool transport = true) {
  auto add = [](const F3& a, const F3& b) { return mkf(a.x + b.x, a.y + b.y, a.z + b.z); };
  auto sub = [](const F3& a, const F3& b) { return mkf(a.x - b.x, a.y - b.y, a.z - b.z); };
  auto scl = [](float s, const F3& a) { return mkf(s * a.x, s * a.y, s * a.z); };
  const F3 pos = tof(T.pos);
  const F3 dir = tof(T.dir);
  const float qop = static_cast<float>(T.qop);
  const float tol = static_cast<float>(o.stepTolerance);
  F3 B1;
  if (const int fs = fieldAtF(F, pos, B1)) return fs;
  const double initialH = stepValue(T) * propDir;
  double hD = initialH;
  float h = static_cast<float>(initialH);
  float err, hThird;
  F3 bend1, bend2, bend3, kick1, c2, c3, kick4, dir2, dir3, dirEnd;
  int nTrials = 0;
  while (true) {
    ++nTrials;
    const float halfHQop = h * qop * 0.5f;
    hThird = h / 3.f;
    const float hQuarter = h * 0.25f;
    bend1 = scl(halfHQop, B1);
    kick1 = crossf(dir, bend1);
    dir2 = add(dir, kick1);
    const F3 pos2 = add(pos, scl(hQuarter, add(dir2, dir)));
    F3 B2;
    if (const int fs = fieldAtF(F, pos2, B2)) return fs;
    bend2 = scl(halfHQop, B2);
    c2 = crossf(dir2, bend2);
    dir3 = add(dir, c2);
    c3 = crossf(dir3, bend2);
    const F3 dir4 = add(dir, c3);
    dirEnd = sub(scl(2.f, dir4), dir);
    const F3 pos3 = add(pos, scl(h, dir4));
    F3 B3;
    if (const int fs = fieldAtF(F, pos3, B3)) return fs;
    bend3 = scl(halfHQop, B3);
    kick4 = crossf(dirEnd, bend3);
    const F3 e = add(sub(sub(kick1, c2), c3), kick4);
    err = 2.f * fabsf(h) * ((fabsf(e.x) + fabsf(e.y)) + fabsf(e.z));
    err = fmaxf(1e-20f, err);
    if (err <= 4.0f * tol) break;
    h *= stepScalingF(tol, err);
    hD = static_cast<double>(h);
    if (fabs(hD) < fabs(o.stepSizeCutOff)) return ACTS_KF_STATUS_STEP_SIZE_STALLED;
    if (static_cast<uint32_t>(nTrials) > o.maxTrials) return ACTS_KF_STATUS_STEP_ADJUSTMENT_FAILED;
  }
  const float m = o.mass;
  const double pD = copysign(static_cast<double>(o.absQ), T.qop) / T.qop;
  const double mpD = m / pD;
  const double dtds = sqrt(1 + mpD * mpD);
  T.time += hD * dtds;
  if (transport) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const F3 seed = mkf(static_cast<float>(T.br[c]), static_cast<float>(T.br[3 + c]),
                          static_cast<float>(T.br[6 + c]));
      const F3 tKick1 = crossf(seed, bend1);
      const F3 tDir2 = add(tKick1, seed);
      const F3 tDir3 = add(seed, crossf(tDir2, bend2));
      const F3 tDir4 = add(seed, crossf(tDir3, bend2));
      const F3 tDirEnd = sub(scl(2.f, tDir4), seed);
      const F3 tKick4 = crossf(tDirEnd, bend3);
      const F3 dp = add(add(tDir2, tDir3), tDir4);
      const F3 dd = add(add(add(tKick1, scl(2.f, tDir3)), tDirEnd), tKick4);
      T.tr[c] += static_cast<double>(hThird * dp.x);
      T.tr[3 + c] += static_cast<double>(hThird * dp.y);
      T.tr[6 + c] += static_cast<double>(hThird * dp.z);
      T.br[c] = static_cast<double>(dd.x / 3.f);
      T.br[3 + c] = static_cast<double>(dd.y / 3.f);
      T.br[6 + c] = static_cast<double>(dd.z / 3.f);
    }
    {
      const float invQop = 1.f / qop;
      const float posW = hThird * invQop, dirW = invQop / 3.f;
      const F3 seed = scl(qop, mkf(static_cast<float>(T.br[2]), static_cast<float>(T.br[5]),
                                   static_cast<float>(T.br[8])));
      const F3 tKick1 = add(kick1, crossf(seed, bend1));
      const F3 tDir2 = add(tKick1, seed);
      const F3 tDir3 = add(add(c2, seed), crossf(tDir2, bend2));
      const F3 tDir4 = add(add(c3, seed), crossf(tDir3, bend2));
      const F3 tDirEnd = sub(scl(2.f, tDir4), seed);
      const F3 tKick4 = add(kick4, crossf(tDirEnd, bend3));
      const F3 dp = add(add(tDir2, tDir3), tDir4);
      const F3 dd = add(add(add(tKick1, scl(2.f, tDir3)), tDirEnd), tKick4);
      T.tr[2] += static_cast<double>(posW * dp.x);
      T.tr[5] += static_cast<double>(posW * dp.y);
      T.tr[8] += static_cast<double>(posW * dp.z);
      T.br[2] = static_cast<double>(dirW * dd.x);
      T.br[5] = static_cast<double>(dirW * dd.y);
      T.br[8] = static_cast<double>(dirW * dd.z);
      T.tr37 += hD * m * m / (pD * pD * T.qop * dtds);
    }
  }
  // position: h dir in double, the bends (h/3)(kick1 + c2 + c3) in float;
  // direction: dir + (kick1 + 2 c2 + 2 c3 + kick4) / 3, renormalised
  const F3 bp = scl(hThird, add(add(kick1, c2), c3));
  T.pos.x += hD * T.dir.x + static_cast<double>(bp.x);
  T.pos.y += hD * T.dir.y + static_cast<double>(bp.y);
  T.pos.z += hD * T.dir.z + static_cast<double>(bp.z);
  const F3 bd = scl(1.f / 3.f, add(add(add(kick1, scl(2.f, c2)), scl(2.f, c3)), kick4));
  const V3 nd = mk(T.dir.x + static_cast<double>(bd.x), T.dir.y + static_cast<double>(bd.y),
                   T.dir.z + static_cast<double>(bd.z));
  const double nn = dot(nd, nd);
  double inv = static_cast<double>(rsqrtf(static_cast<float>(nn)));
  inv = inv * (1.5 - (0.5 * nn) * (inv * inv));
  T.dir = mk(nd.x * inv, nd.y * inv, nd.z * inv);
  if (transport) {
    const float twoOverH = 2.f / h;
    T.der[0] = T.dir.x;
    T.der[1] = T.dir.y;
    T.der[2] = T.dir.z;
    T.der[3] = dtds;
    T.der[4] = static_cast<double>(twoOverH * kick4.x);
    T.der[5] = static_cast<double>(twoOverH * kick4.y);
    T.der[6] = static_cast<double>(twoOverH * kick4.z);
  }
  T.pathAcc += hD;
  ++T.nSteps;
  T.nTrials += nTrials;
  const double nextAccuracy = fabs(hD * static_cast<double>(stepScalingF(tol, err)));
  const double previousAccuracy = fabs(T.cAcc);
  const double initialStepLength = fabs(initialH);
  if (nextAccuracy < initialStepLength || nextAccuracy > previousAccuracy) T.cAcc = nextAccuracy;
  return 0;
}

}  // namespace kf
