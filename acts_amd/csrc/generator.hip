// generator.hip -- synthetic muon events for the batched fitter (host code).
// See include/acts_kf_gen.h for the model.  Truth propagation: adaptive RKN4
// with the stepper's error control (EigenStepper.ipp:192-307), surfaces found
// by straight-line intersection of every not-yet-crossed surface within its
// bounds, mean Bethe energy loss at each material slab (MaterialInteractor
// with covariance transport off, PointwiseMaterialInteraction.hpp:165-205).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/acts_kf_gen.h"
#include "kf_device.hpp"
#include "kf_field_host.hpp"

using namespace kf;

namespace {

// splitmix64 counter-based stream per track
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed, uint64_t stream) {
    s = seed * 0x9E3779B97F4A7C15ull ^ (stream + 0x632BE59BD9B4E019ull);
    next();
  }
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {
    double u1 = uniform();
    while (u1 <= 0.0) u1 = uniform();
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(kTwoPi * u2);
  }
};

struct GenTrack {
  double startParams[6], startCov[36], startPlane[12], truthStart[6];
  std::vector<int32_t> seq;
  std::vector<int32_t> measSurf;
  std::vector<double> measVal, measCov, truthMeas;
  std::vector<int32_t> isOutlier;
};

struct GenWorld {
  std::vector<DevSurface> surf;
  std::vector<DevSlab> slab;
  std::vector<uint32_t> flags;
  DevField F{};
  std::vector<double> grid;  // host copy of a field map (F.v points here)
  float mass, absQ;
  PhysConst pc;
};

bool insideBounds(const DevSurface& s, const V3& pos) {
  const V3 loc = toLocal(s, pos);
  if (s.type == ACTS_KF_SURFACE_CYLINDER) return std::fabs(loc.z) <= s.b1;
  if (s.type == ACTS_KF_SURFACE_DISC) {
    const double r = perp(loc);
    return r >= s.b0 && r <= s.b1;
  }
  return true;
}

// One RKN4 step of at most h (shrunk until the error estimate is accepted),
// with the field evaluated at the three RKN4 points; returns the step taken,
// or NaN if the field lookup failed (the particle left the field map).
double rkStepTruth(V3& pos, V3& dir, double qop, const DevField& F, double h,
                   double& accuracy) {
  V3 B1, B2, B3;
  if (fieldAt(F, pos, B1)) return NAN;
  const V3 c1 = cross(dir, B1);
  const V3 k1 = mk(qop * c1.x, qop * c1.y, qop * c1.z);
  double h2 = 0, err = 0;
  V3 k2, k3, k4;
  for (int trial = 0; trial < 100; ++trial) {
    h2 = h * h;
    const double hh = 0.5 * h;
    const V3 p1 = mk(pos.x + hh * dir.x + h2 * 0.125 * k1.x, pos.y + hh * dir.y + h2 * 0.125 * k1.y,
                     pos.z + hh * dir.z + h2 * 0.125 * k1.z);
    if (fieldAt(F, p1, B2)) return NAN;
    V3 a = cross(mk(dir.x + hh * k1.x, dir.y + hh * k1.y, dir.z + hh * k1.z), B2);
    k2 = mk(qop * a.x, qop * a.y, qop * a.z);
    a = cross(mk(dir.x + hh * k2.x, dir.y + hh * k2.y, dir.z + hh * k2.z), B2);
    k3 = mk(qop * a.x, qop * a.y, qop * a.z);
    const V3 p2 = mk(pos.x + h * dir.x + h2 * 0.5 * k3.x, pos.y + h * dir.y + h2 * 0.5 * k3.y,
                     pos.z + h * dir.z + h2 * 0.5 * k3.z);
    if (fieldAt(F, p2, B3)) return NAN;
    a = cross(mk(dir.x + h * k3.x, dir.y + h * k3.y, dir.z + h * k3.z), B3);
    k4 = mk(qop * a.x, qop * a.y, qop * a.z);
    err = h2 * (std::fabs(k1.x - k2.x - k3.x + k4.x) + std::fabs(k1.y - k2.y - k3.y + k4.y) +
                std::fabs(k1.z - k2.z - k3.z + k4.z));
    err = std::max(1e-20, err);
    if (err <= 4e-4) break;
    h *= std::clamp(std::sqrt(std::sqrt(1e-4 / err)), 0.25, 4.0);
  }
  pos = mk(pos.x + h * dir.x + h2 / 6. * (k1.x + k2.x + k3.x),
           pos.y + h * dir.y + h2 / 6. * (k1.y + k2.y + k3.y),
           pos.z + h * dir.z + h2 / 6. * (k1.z + k2.z + k3.z));
  dir = normalized(mk(dir.x + h / 6. * (k1.x + 2. * (k2.x + k3.x) + k4.x),
                      dir.y + h / 6. * (k1.y + 2. * (k2.y + k3.y) + k4.y),
                      dir.z + h / 6. * (k1.z + 2. * (k2.z + k3.z) + k4.z)));
  const double next = std::fabs(h * std::clamp(std::sqrt(std::sqrt(1e-4 / err)), 0.25, 4.0));
  accuracy = next;
  return h;
}

// curvilinear frame (CurvilinearSurface.cpp:25-39)
void curvilinearPlane(const V3& pos, const V3& T, double* plane12) {
  const bool standard = std::fabs(T.z) < 0.999995;
  const V3 ref = standard ? mk(0, 0, 1) : mk(1, 0, 0);
  const V3 U = normalized(cross(ref, T));
  const V3 V = cross(T, U);
  const double R[9] = {U.x, V.x, T.x, U.y, V.y, T.y, U.z, V.z, T.z};
  std::memcpy(plane12, R, sizeof(R));
  plane12[9] = pos.x;
  plane12[10] = pos.y;
  plane12[11] = pos.z;
}

void generateTrack(const GenWorld& w, const acts_kf_gen_config& cfg, int64_t index,
                   GenTrack& out) {
  const int attempts = std::max(1, cfg.max_attempts);
  for (int attempt = 0; attempt < attempts; ++attempt) {
    Rng rng(cfg.seed, static_cast<uint64_t>(index) * 64 + attempt);
    out.seq.clear();
    out.measSurf.clear();
    out.measVal.clear();
    out.measCov.clear();
    out.truthMeas.clear();
    out.isOutlier.clear();
    double pt;
    const double u = rng.uniform();
    if (cfg.pt_log_uniform)
      pt = std::exp(std::log(cfg.pt_min) + u * (std::log(cfg.pt_max) - std::log(cfg.pt_min)));
    else
      pt = cfg.pt_min + u * (cfg.pt_max - cfg.pt_min);
    const double eta = cfg.eta_min + rng.uniform() * (cfg.eta_max - cfg.eta_min);
    const double phi = cfg.phi_min + rng.uniform() * (cfg.phi_max - cfg.phi_min);
    const double vz = cfg.vertex_sigma_z * rng.normal();
    const double q = rng.uniform() < 0.5 ? -1.0 : 1.0;
    const double theta = 2.0 * std::atan(std::exp(-eta));
    const double p = pt / std::sin(theta);
    double qop = q * w.absQ / p;
    V3 pos = mk(0, 0, vz);
    V3 dir = mk(std::cos(phi) * std::sin(theta), std::sin(phi) * std::sin(theta),
                std::cos(theta));
    double time = 0.0;

    // start parameters on the curvilinear plane at the vertex
    curvilinearPlane(pos, dir, out.startPlane);
    out.truthStart[0] = 0;
    out.truthStart[1] = 0;
    out.truthStart[2] = phi;
    out.truthStart[3] = theta;
    out.truthStart[4] = qop;
    out.truthStart[5] = 0;
    const double* ss = cfg.start_sigma;
    double sp[6];
    sp[0] = ss[0] * rng.normal();
    sp[1] = ss[1] * rng.normal();
    sp[2] = phi + ss[2] * rng.normal();
    sp[3] = theta + ss[3] * rng.normal();
    sp[4] = qop * (1.0 + ss[4] * rng.normal());
    sp[5] = ss[5] * rng.normal();
    normalizePhiTheta(sp[2], sp[3]);
    std::memcpy(out.startParams, sp, sizeof(sp));
    const double sig[6] = {ss[0], ss[1], ss[2], ss[3], ss[4] * std::fabs(sp[4]), ss[5]};
    for (int i = 0; i < 36; ++i) out.startCov[i] = 0;
    for (int i = 0; i < 6; ++i) out.startCov[7 * i] = sig[i] * sig[i];

    // truth propagation through the detector
    const int nS = static_cast<int>(w.surf.size());
    std::vector<char> crossed(nS, 0);
    double accuracy = 10000.0;
    double path = 0.0;
    V3 B0;
    if (fieldAt(w.F, pos, B0)) continue;  // vertex outside the field map
    const double Bn = std::sqrt(B0.x * B0.x + B0.y * B0.y + B0.z * B0.z);
    const double loopLimit = Bn > 0 ? 0.5 * kTwoPi * p / Bn : 1e9;
    int target = -1;
    for (int it = 0; it < 2000 && path < loopLimit; ++it) {
      if (target < 0) {
        double best = INFINITY;
        for (int s = 0; s < nS; ++s) {
          if (crossed[s]) continue;
          const Isect x = intersect(w.surf[s], pos, dir, kOnSurfaceTolerance);
          for (int i = 0; i < x.size; ++i) {
            if (x.status[i] == kUnreachable || x.path[i] <= kOnSurfaceTolerance) continue;
            const double s_ = x.path[i];
            const V3 at = mk(pos.x + s_ * dir.x, pos.y + s_ * dir.y, pos.z + s_ * dir.z);
            if (s_ < best && insideBounds(w.surf[s], at)) {
              best = s_;
              target = s;
            }
            break;
          }
        }
        if (target < 0) break;  // left the detector
      }
      const Isect x = intersect(w.surf[target], pos, dir, kOnSurfaceTolerance);
      double dist = INFINITY;
      for (int i = 0; i < x.size; ++i)
        if (x.status[i] != kUnreachable && x.path[i] > -0.1) {
          dist = x.path[i];
          break;
        }
      if (!std::isfinite(dist)) {
        crossed[target] = 1;
        target = -1;
        continue;
      }
      if (std::fabs(dist) >= kOnSurfaceTolerance) {
        const double h = std::min(dist, accuracy);
        const double taken = rkStepTruth(pos, dir, qop, w.F, h, accuracy);
        if (std::isnan(taken)) break;  // left the field map: the track ends
        const double pp = 1.0 / std::fabs(qop);
        const double mp = w.mass / pp;
        time += taken * std::sqrt(1 + mp * mp);
        path += taken;
        continue;
      }
      // on the target surface
      const int s = target;
      crossed[s] = 1;
      target = -1;
      const DevSurface& S = w.surf[s];
      if (!insideBounds(S, pos)) continue;
      out.seq.push_back(s);
      double l0 = 0, l1 = 0;
      globalToLocal(S, pos, l0, l1);
      if ((w.flags[s] & ACTS_KF_SURFACE_SENSITIVE) && rng.uniform() >= cfg.hole_probability) {
        double s0 = cfg.sigma_loc0, s1 = cfg.sigma_loc1;
        if (S.type == ACTS_KF_SURFACE_DISC) s1 = cfg.sigma_loc1 / l0;  // r*phi -> phi
        out.measSurf.push_back(s);
        out.measVal.push_back(l0 + s0 * rng.normal());
        out.measVal.push_back(l1 + s1 * rng.normal());
        const double mc[4] = {s0 * s0, 0.0, 0.0, s1 * s1};
        out.measCov.insert(out.measCov.end(), mc, mc + 4);
        const double tr[6] = {l0,  l1, std::atan2(dir.y, dir.x), std::atan2(perp(dir), dir.z),
                              qop, time};
        out.truthMeas.insert(out.truthMeas.end(), tr, tr + 6);
        out.isOutlier.push_back(0);
      }
      if (S.mat >= 0) {
        // mean energy loss (MaterialInteractor, covariance transport off)
        DevSlab sl = w.slab[S.mat];
        const float sc = static_cast<float>(pathCorrection(S, pos, dir));
        sl.thickness *= sc;
        sl.tInX0 *= sc;
        sl.tInL0 *= sc;
        const float eLoss =
            energyLossBethe(w.pc, sl, w.mass, static_cast<float>(qop), w.absQ);
        const double mom = 1.0 / std::fabs(qop) * w.absQ;
        const double m = w.mass;
        const double nextE = std::sqrt(m * m + mom * mom) - eLoss;
        double nextP = (m < nextE) ? cathetus(nextE, m) : 0;
        nextP = std::max(0.01, nextP);
        qop = std::copysign(static_cast<double>(w.absQ), qop) / nextP;
      }
      if (1.0 / std::fabs(qop) < 0.05) break;  // stopped
    }
    const int nMeas = static_cast<int>(out.measSurf.size());
    if (nMeas < std::max(1, cfg.min_measurements) && attempt + 1 < attempts) continue;

    // outliers: displace 1..k measurements by outlier_distance sigmas
    if (nMeas > 0 && cfg.outlier_fraction > 0 && rng.uniform() < cfg.outlier_fraction) {
      const int lo = std::max(1, cfg.outlier_min), hi = std::max(lo, cfg.outlier_max);
      int k = lo + static_cast<int>(rng.uniform() * (hi - lo + 1));
      k = std::min(k, nMeas);
      for (int j = 0; j < k; ++j) {
        int m = static_cast<int>(rng.uniform() * nMeas);
        for (int tries = 0; tries < nMeas && out.isOutlier[m]; ++tries) m = (m + 1) % nMeas;
        out.isOutlier[m] = 1;
        out.measVal[2 * m] += cfg.outlier_distance * std::sqrt(out.measCov[4 * m]);
        out.measVal[2 * m + 1] -= cfg.outlier_distance * std::sqrt(out.measCov[4 * m + 3]);
      }
    }
    return;
  }
}

// SolenoidBField (SolenoidBField.cpp:24-162).  The reference passes k^2 to
// Boost's ellint_1/ellint_2, which take the modulus, so the integrals here
// are K and E of modulus k^2.
void ellipticKE(double k, double& K, double& E) {
  // arithmetic-geometric mean (Abramowitz & Stegun 17.6)
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

struct Solenoid {
  double radius, length, dz, R2, scale;
  int nCoils;
  Solenoid(double radius_, double length_, int nCoils_, double bMagCenter)
      : radius(radius_), length(length_), nCoils(nCoils_) {
    dz = length / nCoils;
    R2 = radius * radius;
    double br, bz;
    multiCoil(0.0, 0.0, 1.0, br, bz);
    scale = bMagCenter / std::sqrt(br * br + bz * bz);
  }
  double k2(double r, double z) const {
    return 4 * radius * r / ((radius + r) * (radius + r) + z * z);
  }
  void multiCoil(double r, double z, double sc, double& br, double& bz) const {
    br = 0;
    bz = 0;
    for (int coil = 0; coil < nCoils; ++coil) {
      const double zs = z + length * 0.5 - dz * (coil + 0.5);
      double vr, vz;
      coilField(r, zs, sc, vr, vz);
      br += vr;
      bz += vz;
    }
  }
  // B_r and B_z of one coil, the elliptic integrals shared between the two
  void coilField(double rs, double z, double sc, double& vr, double& vz) const {
    const double r = std::fabs(rs);
    if (r == 0) {
      vr = 0.;
      vz = sc / 2. * R2 / (std::sqrt(R2 + z * z) * (R2 + z * z));
      return;
    }
    const double k_2 = k2(r, z);
    const double k = std::sqrt(k_2);
    double K, E;
    ellipticKE(k_2, K, E);
    const double cr = sc * k * z / (4 * kPi * std::sqrt(radius * r * r * r));
    const double Bsr = (2. - k_2) / (2. - 2. * k_2) * E - K;
    vr = r / rs * cr * Bsr;
    const double cz = sc * k / (4 * kPi * std::sqrt(radius * r));
    const double Bsz = ((radius + r) * k_2 - 2. * r) / (2. * r * (1. - k_2)) * E + K;
    vz = cz * Bsz;
  }
  void rzField(double r, double z, double& br, double& bz) const {
    multiCoil(r, z, scale, br, bz);
  }
  // getField: (0, 0, B_z) plus B_r along the normalised (x, y, 0)
  void compose(double x, double y, double r, double br, double bz, double* out) const {
    out[0] = 0;
    out[1] = 0;
    out[2] = bz;
    if (r != 0.) {
      const double n = std::sqrt((x * x + y * y) + 0.0 * 0.0);
      out[0] = out[0] + (x / n) * br;
      out[1] = out[1] + (y / n) * br;
      out[2] = out[2] + (0.0 / n) * br;
    }
  }
};

}  // namespace

struct acts_kf_gen {
  std::vector<GenTrack> tracks;
  int64_t nSeq = 0, nMeas = 0;
};

extern "C" {

int acts_kf_gen_run(const acts_kf_gen_config* cfg, const acts_kf_surface* surfaces,
                    int32_t nSurfaces, const acts_kf_material_slab* materials,
                    int32_t nMaterials, const acts_kf_field* field,
                    const acts_kf_options* particle, acts_kf_gen** out) {
  if (!cfg || !surfaces || nSurfaces <= 0 || !field || !particle || !out || cfg->n_tracks < 0)
    return ACTS_KF_ERR_INVALID_ARGUMENT;
  GenWorld w;
  if (field->kind == ACTS_KF_FIELD_GRID_XYZ) {
    if (!buildGridField(*field, w.F, w.grid)) return ACTS_KF_ERR_INVALID_ARGUMENT;
    w.F.v = w.grid.data();
  } else if (field->kind == ACTS_KF_FIELD_CONSTANT) {
    w.F.kind = ACTS_KF_FIELD_CONSTANT;
    for (int i = 0; i < 3; ++i) w.F.B[i] = field->constant[i];
  } else {
    return ACTS_KF_ERR_UNSUPPORTED;
  }
  w.surf.resize(nSurfaces);
  w.flags.resize(nSurfaces);
  for (int i = 0; i < nSurfaces; ++i) {
    DevSurface& d = w.surf[i];
    std::memcpy(d.R, surfaces[i].rotation, sizeof(d.R));
    std::memcpy(d.t, surfaces[i].translation, sizeof(d.t));
    invert3(d.R, d.Ri);
    for (int r = 0; r < 3; ++r)
      d.ti[r] = -((d.Ri[3 * r] * d.t[0] + d.Ri[3 * r + 1] * d.t[1]) + d.Ri[3 * r + 2] * d.t[2]);
    d.b0 = surfaces[i].bounds[0];
    d.b1 = surfaces[i].bounds[1];
    d.type = surfaces[i].type;
    d.flags = static_cast<int32_t>(surfaces[i].flags);
    d.mat = (surfaces[i].material >= 0 && surfaces[i].material < nMaterials)
                ? surfaces[i].material
                : -1;
    w.flags[i] = surfaces[i].flags;
  }
  w.slab.resize(std::max(nMaterials, 1));
  for (int i = 0; i < nMaterials; ++i) {
    const acts_kf_material_slab& m = materials[i];
    DevSlab& d = w.slab[i];
    d.thickness = m.thickness;
    d.tInX0 = m.thickness / m.x0;
    d.tInL0 = m.thickness / m.l0;
    d.molarElectronRho = m.z * m.molar_rho;
    d.meanI = static_cast<float>(static_cast<double>(1e-9 * 16.0L) * std::pow(m.z, 0.9f));
    d.vacuum = m.ar <= 0.f;
  }
  w.mass = particle->mass;
  w.absQ = particle->abs_charge;
  w.pc.me = static_cast<float>(static_cast<double>(1e-3 * 0.5109989461L));
  w.pc.k = static_cast<float>(static_cast<double>(1e-3 * 30.7075L));
  w.pc.plasmaScale = static_cast<float>(static_cast<double>(1e-9 * 28.816L));
  w.pc.highland = static_cast<double>(1e-3 * 13.6L);
  w.pc.rossiGreisen = static_cast<double>(1e-3 * 17.5L);
  w.pc.minP = 0.01;

  acts_kf_gen* g = new acts_kf_gen();
  g->tracks.resize(cfg->n_tracks);
  int nThreads = cfg->n_threads > 0 ? cfg->n_threads
                                    : static_cast<int>(std::thread::hardware_concurrency());
  nThreads = std::max(1, std::min(nThreads, 64));
  std::atomic<int64_t> next{0};
  auto worker = [&]() {
    while (true) {
      const int64_t t0 = next.fetch_add(256);
      if (t0 >= cfg->n_tracks) break;
      const int64_t t1 = std::min<int64_t>(cfg->n_tracks, t0 + 256);
      for (int64_t t = t0; t < t1; ++t) generateTrack(w, *cfg, t, g->tracks[t]);
    }
  };
  std::vector<std::thread> pool;
  for (int i = 0; i < nThreads; ++i) pool.emplace_back(worker);
  for (auto& th : pool) th.join();
  for (const GenTrack& t : g->tracks) {
    g->nSeq += static_cast<int64_t>(t.seq.size());
    g->nMeas += static_cast<int64_t>(t.measSurf.size());
  }
  *out = g;
  return ACTS_KF_OK;
}

void acts_kf_gen_sizes(const acts_kf_gen* g, int64_t* nTracks, int64_t* nSeq, int64_t* nMeas) {
  if (nTracks) *nTracks = g ? static_cast<int64_t>(g->tracks.size()) : 0;
  if (nSeq) *nSeq = g ? g->nSeq : 0;
  if (nMeas) *nMeas = g ? g->nMeas : 0;
}

void acts_kf_gen_fetch(const acts_kf_gen* g, double* startParams, double* startCov,
                       double* startPlane, int64_t* seqOffsets, int32_t* seqSurfaces,
                       int64_t* measOffsets, int32_t* measSurfaces, double* measValues,
                       double* measCov, double* truthStart, double* truthMeas,
                       int32_t* measIsOutlier) {
  if (!g) return;
  int64_t so = 0, mo = 0;
  const int64_t n = static_cast<int64_t>(g->tracks.size());
  for (int64_t t = 0; t < n; ++t) {
    const GenTrack& tr = g->tracks[t];
    if (startParams) std::memcpy(startParams + 6 * t, tr.startParams, sizeof(tr.startParams));
    if (startCov) std::memcpy(startCov + 36 * t, tr.startCov, sizeof(tr.startCov));
    if (startPlane) std::memcpy(startPlane + 12 * t, tr.startPlane, sizeof(tr.startPlane));
    if (truthStart) std::memcpy(truthStart + 6 * t, tr.truthStart, sizeof(tr.truthStart));
    if (seqOffsets) seqOffsets[t] = so;
    if (measOffsets) measOffsets[t] = mo;
    const size_t ns = tr.seq.size(), nm = tr.measSurf.size();
    if (seqSurfaces && ns) std::memcpy(seqSurfaces + so, tr.seq.data(), ns * sizeof(int32_t));
    if (measSurfaces && nm)
      std::memcpy(measSurfaces + mo, tr.measSurf.data(), nm * sizeof(int32_t));
    if (measValues && nm) std::memcpy(measValues + 2 * mo, tr.measVal.data(), nm * 2 * sizeof(double));
    if (measCov && nm) std::memcpy(measCov + 4 * mo, tr.measCov.data(), nm * 4 * sizeof(double));
    if (truthMeas && nm) std::memcpy(truthMeas + 6 * mo, tr.truthMeas.data(), nm * 6 * sizeof(double));
    if (measIsOutlier && nm)
      std::memcpy(measIsOutlier + mo, tr.isOutlier.data(), nm * sizeof(int32_t));
    so += static_cast<int64_t>(ns);
    mo += static_cast<int64_t>(nm);
  }
  if (seqOffsets) seqOffsets[n] = so;
  if (measOffsets) measOffsets[n] = mo;
}

void acts_kf_gen_free(acts_kf_gen* g) { delete g; }

int acts_kf_solenoid_map(const acts_kf_solenoid* cfg, const int32_t n[3],
                         const double pmin[3], const double pmax[3], double* values,
                         int32_t nThreads) {
  if (!cfg || !n || !pmin || !pmax || !values || cfg->n_coils <= 0 || !(cfg->radius > 0) ||
      !(cfg->length > 0))
    return ACTS_KF_ERR_INVALID_ARGUMENT;
  for (int a = 0; a < 3; ++a)
    if (n[a] < 2 || !(pmax[a] > pmin[a])) return ACTS_KF_ERR_INVALID_ARGUMENT;
  const Solenoid sol(cfg->radius, cfg->length, cfg->n_coils, cfg->b_mag_center);
  std::vector<double> xs[3];
  for (int a = 0; a < 3; ++a) {
    const double step = (pmax[a] - pmin[a]) / static_cast<double>(n[a] - 1);
    xs[a].resize(n[a]);
    for (int i = 0; i < n[a]; ++i) xs[a][i] = pmin[a] + static_cast<double>(i) * step;
  }
  // The (B_r, B_z) part depends on (r, z) only: evaluate it once per distinct
  // transverse radius of the grid.
  std::vector<double> rOf(static_cast<size_t>(n[0]) * n[1]);
  for (int i = 0; i < n[0]; ++i)
    for (int j = 0; j < n[1]; ++j) rOf[i * n[1] + j] = std::sqrt(xs[0][i] * xs[0][i] + xs[1][j] * xs[1][j]);
  std::vector<double> radii(rOf);
  std::sort(radii.begin(), radii.end());
  radii.erase(std::unique(radii.begin(), radii.end()), radii.end());
  const int64_t nr = static_cast<int64_t>(radii.size()), nz = n[2];
  std::vector<double> table(static_cast<size_t>(2 * nr * nz));
  int nt = nThreads > 0 ? nThreads : static_cast<int>(std::thread::hardware_concurrency());
  nt = std::max(1, std::min(nt, 64));
  std::atomic<int64_t> next{0};
  auto work = [&]() {
    for (int64_t r = next++; r < nr; r = next++)
      for (int64_t k = 0; k < nz; ++k)
        sol.rzField(radii[r], xs[2][k], table[2 * (r * nz + k)], table[2 * (r * nz + k) + 1]);
  };
  std::vector<std::thread> pool;
  for (int i = 0; i < nt; ++i) pool.emplace_back(work);
  for (auto& th : pool) th.join();
  std::atomic<int64_t> nextRow{0};
  auto fill = [&]() {
    for (int64_t i = nextRow++; i < n[0]; i = nextRow++)
      for (int j = 0; j < n[1]; ++j) {
        const double r = rOf[i * n[1] + j];
        const int64_t ri = std::lower_bound(radii.begin(), radii.end(), r) - radii.begin();
        for (int k = 0; k < n[2]; ++k) {
          const double br = table[2 * (ri * nz + k)], bz = table[2 * (ri * nz + k) + 1];
          double* out = values + 3 * ((i * n[1] + j) * n[2] + k);
          sol.compose(xs[0][i], xs[1][j], r, br, bz, out);
        }
      }
  };
  pool.clear();
  for (int i = 0; i < nt; ++i) pool.emplace_back(fill);
  for (auto& th : pool) th.join();
  return ACTS_KF_OK;
}

}  // extern "C"
