// kf_field_host.hpp -- host-side set-up of the field description used by the
// kernels (and by the host-side event generator).
#pragma once

#include <cmath>
#include <cstring>
#include <vector>

#include "kf_common.hpp"

namespace kf {

// The reciprocals that let a lookup multiply instead of divide (DevField::
// rw, den, rden, recip): on when, per axis, every bin's upper minus lower
// edge (computed as gridLocate does) is one value and all of them are
// normal, positive numbers.  Divisions are correctly rounded here and on the
// device, so divRecip then returns what the division would.
inline void setGridReciprocals(DevField& F) {
  bool ok = true;
  for (int a = 0; a < 3; ++a) {
    const double w = F.width[a];
    ok = ok && std::isnormal(w) && w > 0;
    double den = 0;
    for (int32_t b = 1; ok && b <= F.nb[a]; ++b) {
      const double ll = F.min[a] + static_cast<double>(b - 1) * w;
      const double ur = F.min[a] + static_cast<double>(b) * w;
      const double d = ur - ll;
      if (b == 1) den = d;
      ok = d == den;
    }
    ok = ok && std::isnormal(den) && den > 0;
    F.rw[a] = ok ? 1.0 / w : 0.0;
    F.den[a] = ok ? den : 0.0;
    F.rden[a] = ok ? 1.0 / den : 0.0;
  }
  F.recip = ok ? 1 : 0;
}

// Axes as fieldMapXYZ builds them from the grid points (BFieldMapUtils.cpp:
// 109-186, Helpers.hpp:289-310, Axis.hpp:45), and the point values copied
// into an array with one extra, zero, point per axis (the overflow bin).
// `F.v` is left for the caller to point at the (device or host) copy of
// `padded`.  Returns false for an invalid grid.
inline bool buildGridField(const acts_kf_field& in, DevField& F, std::vector<double>& padded) {
  if (!in.values) return false;
  for (int a = 0; a < 3; ++a)
    if (in.n[a] < 2 || in.n[a] > (1 << 20) || !(in.point_max[a] > in.point_min[a])) return false;
  F.kind = ACTS_KF_FIELD_GRID_XYZ;
  for (int a = 0; a < 3; ++a) {
    const int32_t n = in.n[a];
    const double step = (in.point_max[a] - in.point_min[a]) / static_cast<double>(n - 1);
    const double axisMax = in.point_max[a] + step;
    F.nb[a] = n;
    F.min[a] = in.point_min[a];
    F.width[a] = (axisMax - F.min[a]) / static_cast<double>(n);
    F.upper[a] = F.min[a] + static_cast<double>(n - 1) * F.width[a];
  }
  setGridReciprocals(F);
  const int64_t n0 = in.n[0], n1 = in.n[1], n2 = in.n[2];
  const int64_t p1 = n1 + 1, p2 = n2 + 1;
  padded.assign(static_cast<size_t>(3 * (n0 + 1) * p1 * p2), 0.0);
  for (int64_t i = 0; i < n0; ++i)
    for (int64_t j = 0; j < n1; ++j)
      std::memcpy(&padded[3 * ((i * p1 + j) * p2)], in.values + 3 * ((i * n1 + j) * n2),
                  sizeof(double) * 3 * n2);
  F.v = nullptr;
  return true;
}

}  // namespace kf
