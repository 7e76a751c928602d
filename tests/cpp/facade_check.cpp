// facade_check.cpp -- the C++ facade (include/acts_kf.hpp) used the way the
// reference's KalmanFitterFunctionImpl calls KalmanFitter::fit
// (Examples/Algorithms/TrackFitting/src/KalmanFitterFunction.cpp:150-157):
// one call per track, and the batched form.  Checks that both give the
// same results bit for bit and that the fits succeed, and that the
// Navigator overload (no surface sequence) gives the same fits.
//
// Usage: facade_check [n_tracks]   (exit status 0 = pass)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "acts_kf.hpp"
#include "acts_kf_gen.h"

namespace {

constexpr double kTesla = 0.000299792458;
constexpr double kPi = 3.14159265358979323846;

std::vector<acts_kf_surface> barrel() {
  const double radii[] = {32, 72, 116, 172, 260, 360, 500, 660, 820, 1000};
  std::vector<acts_kf_surface> s;
  for (double r : radii) {
    acts_kf_surface x;
    std::memset(&x, 0, sizeof(x));
    x.type = ACTS_KF_SURFACE_CYLINDER;
    x.material = 0;
    x.flags = ACTS_KF_SURFACE_SENSITIVE;
    x.rotation[0] = x.rotation[4] = x.rotation[8] = 1.0;
    x.bounds[0] = r;
    x.bounds[1] = 3000.0;
    s.push_back(x);
  }
  return s;
}

}  // namespace

int main(int argc, char** argv) {
  const int64_t n = argc > 1 ? std::atoll(argv[1]) : 200;
  const std::vector<acts_kf_surface> surfaces = barrel();
  // 0.3 mm silicon (X0 93.70 mm, L0 465.2 mm, Ar 28.0855, Z 14, 2.329 g/cm3)
  std::vector<acts_kf_material_slab> mats(1);
  mats[0].x0 = 93.70f;
  mats[0].l0 = 465.2f;
  mats[0].ar = 28.0855f;
  mats[0].z = 14.f;
  mats[0].molar_rho = static_cast<float>((2.329 / 28.0855) / 1000.0);
  mats[0].thickness = 0.3f;
  acts_kf_field field;
  std::memset(&field, 0, sizeof(field));
  field.kind = ACTS_KF_FIELD_CONSTANT;
  field.constant[2] = 2.0 * kTesla;
  const acts_kf_options opts = acts_amd::KalmanFitter::defaultOptions();

  acts_kf_gen_config cfg;
  std::memset(&cfg, 0, sizeof(cfg));
  cfg.n_tracks = n;
  cfg.seed = 99;
  cfg.pt_min = 1.0;
  cfg.pt_max = 10.0;
  cfg.eta_min = -1.0;
  cfg.eta_max = 1.0;
  cfg.phi_min = -kPi;
  cfg.phi_max = kPi;
  cfg.vertex_sigma_z = 50.0;
  cfg.sigma_loc0 = 0.015;
  cfg.sigma_loc1 = 0.05;
  const double start[6] = {0.15, 0.5, 1e-3, 1e-3, 0.1, 299.792458};
  for (int i = 0; i < 6; ++i) cfg.start_sigma[i] = start[i];
  cfg.min_measurements = 1;
  cfg.max_attempts = 16;
  acts_kf_gen* gen = nullptr;
  if (acts_kf_gen_run(&cfg, surfaces.data(), static_cast<int32_t>(surfaces.size()),
                      mats.data(), 1, &field, &opts, &gen) != ACTS_KF_OK) {
    std::fprintf(stderr, "generator failed: %s\n", acts_kf_last_error());
    return 2;
  }
  int64_t nt = 0, nseq = 0, nmeas = 0;
  acts_kf_gen_sizes(gen, &nt, &nseq, &nmeas);
  std::vector<double> sp(6 * nt), sc(36 * nt), pl(12 * nt), mv(2 * nmeas), mc(4 * nmeas);
  std::vector<int64_t> so(nt + 1), mo(nt + 1);
  std::vector<int32_t> ss(nseq), ms(nmeas);
  acts_kf_gen_fetch(gen, sp.data(), sc.data(), pl.data(), so.data(), ss.data(), mo.data(),
                    ms.data(), mv.data(), mc.data(), nullptr, nullptr, nullptr);
  acts_kf_gen_free(gen);

  // per-track inputs in the facade's types
  std::vector<acts_amd::BoundTrackParameters> starts(nt);
  std::vector<std::vector<acts_amd::Measurement>> meas(nt);
  std::vector<std::vector<int32_t>> seqs(nt);
  for (int64_t t = 0; t < nt; ++t) {
    for (int i = 0; i < 6; ++i) starts[t].params[i] = sp[6 * t + i];
    for (int i = 0; i < 36; ++i) starts[t].covariance[i] = sc[36 * t + i];
    for (int i = 0; i < 9; ++i) starts[t].planeRotation[i] = pl[12 * t + i];
    for (int i = 0; i < 3; ++i) starts[t].planeCenter[i] = pl[12 * t + 9 + i];
    seqs[t].assign(ss.begin() + so[t], ss.begin() + so[t + 1]);
    for (int64_t m = mo[t]; m < mo[t + 1]; ++m) {
      acts_amd::Measurement x;
      x.surface = ms[m];
      x.values = {mv[2 * m], mv[2 * m + 1]};
      x.covariance = {mc[4 * m], mc[4 * m + 1], mc[4 * m + 2], mc[4 * m + 3]};
      meas[t].push_back(x);
    }
  }

  try {
    // as the Examples configure the fit (TrackFittingAlgorithm.cpp:89-90,
    // KalmanFitterFunction.cpp:114-115): fitted parameters at a perigee
    // surface at the origin, extrapolated from the first measurement
    acts_kf_options fitOpts = opts;
    fitOpts.reference_surface = ACTS_KF_REFERENCE_PERIGEE;
    fitOpts.reference_strategy = ACTS_KF_EXTRAPOLATE_FIRST;
    acts_amd::KalmanFitter fitter(surfaces, mats, field, fitOpts, 0);
    for (int64_t t = 0; t < nt; ++t)
      fitter.enqueue(meas[t].begin(), meas[t].end(), starts[t], seqs[t]);
    const std::vector<acts_amd::FitResult> batch = fitter.fitBatch();
    int64_t nOk = 0, mismatches = 0;
    for (const auto& r : batch) nOk += r.ok() ? 1 : 0;
    const int64_t nSingle = nt < 8 ? nt : 8;
    for (int64_t t = 0; t < nSingle; ++t) {
      const acts_amd::FitResult one =
          fitter.fit(meas[t].begin(), meas[t].end(), starts[t], seqs[t]);
      const acts_amd::FitResult& b = batch[t];
      bool same = one.status == b.status && one.states.size() == b.states.size() &&
                  one.chi2 == b.chi2 && one.ndf == b.ndf && one.hasReference == b.hasReference &&
                  (!b.ok() || b.hasReference) &&
                  std::memcmp(one.referenceParameters.data(), b.referenceParameters.data(),
                              6 * sizeof(double)) == 0 &&
                  std::memcmp(one.referenceCovariance.data(), b.referenceCovariance.data(),
                              36 * sizeof(double)) == 0;
      for (size_t k = 0; same && k < one.states.size(); ++k) {
        same = one.states[k].surface == b.states[k].surface &&
               one.states[k].flags == b.states[k].flags &&
               std::memcmp(one.states[k].smoothed.data(), b.states[k].smoothed.data(),
                           6 * sizeof(double)) == 0 &&
               std::memcmp(one.states[k].smoothedCovariance.data(),
                           b.states[k].smoothedCovariance.data(), 36 * sizeof(double)) == 0;
      }
      if (!same) ++mismatches;
    }
    // the Navigator overload (no surface sequence): on this barrel the
    // layer navigator finds the generator's surfaces, so the fits are the
    // same bit for bit
    acts_kf_options layerOpts = fitOpts;
    layerOpts.navigator = ACTS_KF_NAVIGATOR_LAYERS;
    acts_amd::KalmanFitter layers(surfaces, mats, field, layerOpts, 0);
    int64_t layerMismatches = 0;
    for (int64_t t = 0; t < nSingle; ++t) {
      const acts_amd::FitResult one = layers.fit(meas[t].begin(), meas[t].end(), starts[t]);
      const acts_amd::FitResult& b = batch[t];
      bool same = one.status == b.status && one.states.size() == b.states.size() &&
                  one.chi2 == b.chi2 &&
                  std::memcmp(one.referenceParameters.data(), b.referenceParameters.data(),
                              6 * sizeof(double)) == 0;
      for (size_t k = 0; same && k < one.states.size(); ++k)
        same = one.states[k].surface == b.states[k].surface &&
               std::memcmp(one.states[k].smoothed.data(), b.states[k].smoothed.data(),
                           6 * sizeof(double)) == 0;
      if (!same) ++layerMismatches;
    }
    std::printf("facade_check: tracks=%lld ok=%lld single_vs_batch_mismatches=%lld "
                "layer_vs_direct_mismatches=%lld\n",
                static_cast<long long>(nt), static_cast<long long>(nOk),
                static_cast<long long>(mismatches), static_cast<long long>(layerMismatches));
    if (mismatches != 0 || layerMismatches != 0 || nOk < nt * 95 / 100) return 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "facade_check: %s\n", e.what());
    return 2;
  }
  return 0;
}
