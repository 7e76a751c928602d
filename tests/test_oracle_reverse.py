"""Pins of the oracle's reverse filtering (KalmanFitter.hpp:884-945): after
the forward filter the track is filtered again backwards from the last
measurement state's filtered parameters (covariance scaled up), through the
same surface sequence, and the first measurement state takes that filter's
filtered parameters as its smoothed parameters.

What fixes the result:
- Without process noise (material effects off) a backward filter whose
  prior carries no information ends, at the first measurement, with the
  information of every measurement, as the smoothed state does, except for
  the forward filter's start parameters, which only the smoothed state
  has.  With the start covariance scaled by 1e3 (its information 1e-3 of
  the generator's) and the reverse covariance scaling at 1e8, the reverse
  result equals the RTS-smoothed first state up to the two filters'
  different linearisation points: measured within 2.5e-3 of the smoothed
  errors and 3e-4 on the covariance for straight tracks through the tilted
  telescope planes, 0.06 and 3e-3 for helices in 2 T.  A dropped update
  (e.g. the start surface's own measurement) loses a sixth of the
  telescope's information and fails the covariance check.
- determineMaterialUpdateMode (PointwiseMaterialInteraction.cpp:16-29)
  allows only post-update material on the start surface, and post-update
  material in the backward direction has factor 0 (ISurfaceMaterial.cpp:
  18-31): the reverse filter never applies the last measurement surface's
  material, so removing it changes nothing, bit for bit, while removing
  the first plane's material (applied before its update going backwards)
  changes the reverse result.
- The forward pass, the summary and the last measurement state's smoothed
  (= its filtered) parameters are the plain fit's, bit for bit; no other
  state has smoothed parameters (NaN); the momentum threshold selects
  exactly the tracks with |1 / filtered q/p| at or below it.
- With material on, the reverse-filtered first state is unbiased against
  the truth with the smoothed state's pull widths.
"""
import dataclasses

import numpy as np
import pytest

import acts_amd as A
from acts_amd import capi
from acts_amd.detector import Detector
from acts_amd.fitter import ALL_STATE_OUTPUTS


def _fit(oracle_mod, det, field, ev, **kw):
    o = A.KalmanFitterOptions(**kw)
    r, _ = oracle_mod.fit(det, field, o, ev, n_threads=0, outputs=ALL_STATE_OUTPUTS)
    return r


def _meas_range(res, ev, t):
    """(first, last) measurement state slots of track t (not outliers)."""
    s0 = ev.seq_offsets[t]
    f = res.state_flags[s0:s0 + res.n_states[t]]
    ok = np.nonzero((f & capi.STATE_HAS_MEASUREMENT) & ~(f & capi.STATE_IS_OUTLIER))[0]
    return s0 + ok[0], s0 + ok[-1]


def _cov(res, slot):
    return res.smoothed_cov[slot].reshape(6, 6)


@pytest.mark.parametrize("workload,n,seed,max_par,max_cov", [
    ("telescope", 600, 11, 1e-2, 2e-3), ("barrel", 400, 12, 0.1, 1e-2)])
def test_reverse_equals_smoothed_without_process_noise(oracle_mod, workload, n, seed,
                                                       max_par, max_cov):
    det, field, ev = A.workload(workload, n_tracks=n, seed=seed)
    ev.start_cov = ev.start_cov * 1e3
    common = dict(multiple_scattering=False, energy_loss=False)
    rts = _fit(oracle_mod, det, field, ev, **common)
    rev = _fit(oracle_mod, det, field, ev, reverse_filtering=True,
               reverse_filtering_cov_scaling=1e8, **common)
    ok = (rts.status == 0) & (rev.status == 0)
    assert np.count_nonzero(ok) >= 0.98 * n
    assert np.all(rev.reversed[ok] == 1) and np.all(rts.reversed[ok] == 0)
    # loc0, loc1, phi, theta (q/p is not measured without a field, and the
    # time never is)
    worst_par, worst_cov = 0.0, 0.0
    for t in np.nonzero(ok)[0]:
        first, _ = _meas_range(rts, ev, t)
        C = _cov(rts, first)[:4, :4]
        sig = np.sqrt(np.diag(C))
        d = (rev.smoothed[first] - rts.smoothed[first])[:4] / sig
        worst_par = max(worst_par, float(np.max(np.abs(d))))
        Cr = _cov(rev, first)[:4, :4]
        rel = np.abs(Cr - C) / np.sqrt(np.outer(np.diag(C), np.diag(C)))
        worst_cov = max(worst_cov, float(np.max(rel)))
    assert worst_par < max_par, worst_par
    assert worst_cov < max_cov, worst_cov


def test_reverse_structure_and_forward_pass(oracle_mod):
    det, field, ev = A.workload("endcap", n_tracks=800, seed=13)
    plain = _fit(oracle_mod, det, field, ev)
    rev = _fit(oracle_mod, det, field, ev, reverse_filtering=True)
    ok = (plain.status == 0) & (rev.status == 0)
    assert np.count_nonzero(ok) >= 0.97 * ev.n_tracks
    for name in ("n_states", "chi2", "ndf", "n_measurements", "n_holes", "n_outliers",
                 "n_steps", "n_step_trials"):
        assert np.array_equal(getattr(plain, name)[ok], getattr(rev, name)[ok]), name
    for t in np.nonzero(ok)[0]:
        s0 = ev.seq_offsets[t]
        sl = slice(s0, s0 + plain.n_states[t])
        for name in ("predicted", "predicted_cov", "filtered", "filtered_cov", "jacobian",
                     "state_flags", "state_surface"):
            assert np.array_equal(getattr(plain, name)[sl], getattr(rev, name)[sl]), name
        first, last = _meas_range(plain, ev, t)
        assert np.array_equal(rev.smoothed[last], rev.filtered[last])
        assert np.array_equal(rev.smoothed_cov[last], rev.filtered_cov[last].reshape(-1))
        for k in range(s0, s0 + plain.n_states[t]):
            if k in (first, last):
                assert np.all(np.isfinite(rev.smoothed[k]))
            else:
                assert np.all(np.isnan(rev.smoothed[k])) and np.all(np.isnan(rev.smoothed_cov[k]))


def _without_material(det, index):
    d = Detector(name=det.name)
    d.materials = list(det.materials)
    d.kinds = det.kinds
    d.surfaces = [dataclasses.replace(s, material=-1) if i == index else s
                  for i, s in enumerate(det.surfaces)]
    return d


def test_start_surface_material_is_not_applied(oracle_mod):
    det, field, ev = A.workload("telescope", n_tracks=500, seed=14)
    kw = dict(reverse_filtering=True)
    ref = _fit(oracle_mod, det, field, ev, **kw)
    ok = ref.status == 0
    assert np.count_nonzero(ok) >= 0.98 * ev.n_tracks
    # the last plane is every track's last measurement surface: its material
    # is never applied by the reverse filter (nor after the forward one's
    # last update, before any state)
    last = _fit(oracle_mod, _without_material(det, len(det.surfaces) - 1), field, ev, **kw)
    assert np.array_equal(last.status, ref.status)
    for name in ("smoothed", "smoothed_cov", "filtered", "filtered_cov"):
        a, b = getattr(last, name), getattr(ref, name)
        assert np.array_equal(a, b, equal_nan=True), name
    # the first plane's material enters the reverse filter before its update
    first = _fit(oracle_mod, _without_material(det, 0), field, ev, **kw)
    slots = np.array([_meas_range(ref, ev, t)[0] for t in np.nonzero(ok)[0]])
    assert not np.array_equal(first.smoothed_cov[slots], ref.smoothed_cov[slots])


def test_momentum_threshold_selects_tracks(oracle_mod):
    det, field, ev = A.workload("endcap", n_tracks=600, seed=15)
    threshold = 4.0
    plain = _fit(oracle_mod, det, field, ev)
    thr = _fit(oracle_mod, det, field, ev, reverse_filtering_momentum_threshold=threshold)
    ok = (plain.status == 0) & (thr.status == 0)
    expect = np.zeros(ev.n_tracks, bool)
    for t in np.nonzero(ok)[0]:
        _, last = _meas_range(plain, ev, t)
        expect[t] = abs(1.0 / plain.filtered[last][4]) <= threshold
    assert 0 < np.count_nonzero(expect) < np.count_nonzero(ok)
    assert np.array_equal(thr.reversed[ok] == 1, expect[ok])
    # tracks above the threshold are smoothed as usual
    keep = ok & ~expect
    for t in np.nonzero(keep)[0][:100]:
        s0 = ev.seq_offsets[t]
        sl = slice(s0, s0 + plain.n_states[t])
        assert np.array_equal(plain.smoothed[sl], thr.smoothed[sl])


def test_reverse_first_state_pulls(oracle_mod):
    """Pulls of the first measurement state against the generator's truth
    there: unbiased, with widths close to the RTS-smoothed state's (the
    truth has no random scattering, so neither is exactly 1)."""
    det, field, ev = A.workload("barrel", n_tracks=1500, seed=16)
    rts = _fit(oracle_mod, det, field, ev)
    rev = _fit(oracle_mod, det, field, ev, reverse_filtering=True)
    ok = (rev.status == 0) & (rts.status == 0)
    assert np.count_nonzero(ok) >= 0.98 * ev.n_tracks
    pulls = {"rts": [], "rev": []}
    for t in np.nonzero(ok)[0]:
        first, _ = _meas_range(rev, ev, t)
        m = rev.state_measurement[first]
        truth = ev.truth_meas[m]
        for name, r in (("rts", rts), ("rev", rev)):
            C = _cov(r, first)
            pulls[name].append((r.smoothed[first][:2] - truth[:2]) / np.sqrt(np.diag(C)[:2]))
    p = {k: np.array(v) for k, v in pulls.items()}
    assert len(p["rev"]) > 1000
    assert np.all(np.abs(p["rev"].mean(axis=0)) < 0.1)
    assert np.all(np.abs(p["rev"].std(axis=0) - p["rts"].std(axis=0)) < 0.08)
