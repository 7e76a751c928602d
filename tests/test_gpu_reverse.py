"""GPU parity of reverse filtering (KalmanFitter.hpp:884-945): the HIP
fitter with reverse_filtering (or the Examples' momentum threshold) against
the CPU oracle on identical inputs, to the FP64 bar: identical status,
`reversed` flags and state sequences; predicted/filtered parameters,
covariances and jacobians within 1e-10; the smoothed parameters of the
first measurement state (the reverse filter's) and of the last (its own
filtered ones) within 1e-10, their covariances within 5e-8; NaN in the same
smoothed slots (every other state of a reversed track); the perigee
extrapolated from the reverse-filtered first state to the same bar.
Workloads: barrel, endcap and field map, the telescope of the reference's
fitter tests (1D strips, zero field), the ragged batch with holes and
outliers; the MbfSmoother and SympyStepper options; mixed precision within
its bounds of the FP64 oracle.
"""
import numpy as np
import pytest

import acts_amd as A
from acts_amd import capi
from acts_amd.events import WORKLOADS
from acts_amd.fitter import ALL_STATE_OUTPUTS

from parity import compare, cov_rel_err, error_report, param_rel_err

pytestmark = pytest.mark.gpu

OUTPUTS = ALL_STATE_OUTPUTS + ("reference_params", "reference_cov")


def _fit(det, field, ev, **kw):
    fitter = A.KalmanFitter(det, field, A.KalmanFitterOptions(**kw))
    try:
        return fitter.fit(ev, outputs=OUTPUTS)
    finally:
        fitter.close()


def _check(got, ref, ev, n):
    """NaN slots identical, then the FP64 bar on every finite value."""
    assert np.array_equal(got.reversed, ref.reversed)
    ok = (got.status == 0) & (ref.status == 0)
    so = ev.seq_offsets
    idx = np.concatenate([np.arange(so[t], so[t] + ref.n_states[t]) for t in np.nonzero(ok)[0]])
    for name in ("smoothed", "smoothed_cov"):
        g, r = getattr(got, name), getattr(ref, name)
        assert np.array_equal(np.isnan(g[idx]), np.isnan(r[idx])), name
        g[np.isnan(g)] = 0.0
        r[np.isnan(r)] = 0.0
    rep = compare(got, ref, ev, 1e-10)
    print(rep["max_rel_err"], {k: rep[k] for k in ("n_ok", "status_mismatch")}, flush=True)
    assert rep["ok"], rep
    assert np.count_nonzero(ok) >= 0.95 * n
    return ok


@pytest.mark.parametrize("workload,n,seed,extra", [
    ("barrel", 1000, 61, {}),
    ("endcap", 3000, 62, {}),
    ("solenoid", 2000, 63, {}),
    ("telescope", 2000, 64, {}),
    ("ragged", 3000, 65, {}),
    ("endcap", 2000, 66, dict(smoother=capi.SMOOTHER_MBF)),
    ("endcap", 2000, 67, dict(stepper=capi.STEPPER_SYMPY)),
])
def test_reverse_filtering_parity(oracle_mod, workload, n, seed, extra):
    det, field, ev = A.workload(workload, n_tracks=n, seed=seed)
    kw = dict(WORKLOADS[workload].get("options", {}))
    kw.update(extra)
    kw.update(reverse_filtering=True)
    print("reverse", workload, n, seed, extra, flush=True)
    got = _fit(det, field, ev, **kw)
    ref, _ = oracle_mod.fit(det, field, A.KalmanFitterOptions(**kw), ev, n_threads=0,
                            outputs=OUTPUTS)
    ok = _check(got, ref, ev, n)
    assert np.all(got.reversed[ok] == 1)


def test_reverse_threshold_and_perigee(oracle_mod):
    """The Examples' configuration with its reverse-filtering logic switched
    on: tracks at or below 4 GeV at their last measurement are filtered in
    reverse, the others smoothed; the perigee from the first measurement."""
    det, field, ev = A.workload("endcap", n_tracks=3000, seed=68)
    kw = dict(reverse_filtering_momentum_threshold=4.0, smoother=capi.SMOOTHER_MBF,
              stepper=capi.STEPPER_SYMPY, reference_surface=capi.REFERENCE_PERIGEE,
              reference_strategy=capi.EXTRAPOLATE_FIRST)
    got = _fit(det, field, ev, **kw)
    ref, _ = oracle_mod.fit(det, field, A.KalmanFitterOptions(**kw), ev, n_threads=0,
                            outputs=OUTPUTS)
    ok = _check(got, ref, ev, ev.n_tracks)
    nrev = int(np.count_nonzero(got.reversed[ok]))
    assert 0.1 * np.count_nonzero(ok) < nrev < 0.9 * np.count_nonzero(ok)
    assert param_rel_err(got.reference_params[ok], ref.reference_params[ok]).max() <= 1e-10
    assert cov_rel_err(got.reference_cov[ok], ref.reference_cov[ok]).max() <= 5e-8


def test_reverse_filtering_mixed_precision(oracle_mod):
    det, field, ev = A.workload("solenoid", n_tracks=3000, seed=69)
    got = _fit(det, field, ev, reverse_filtering=True, precision=capi.PRECISION_MIXED)
    ref, _ = oracle_mod.fit(det, field, A.KalmanFitterOptions(reverse_filtering=True), ev,
                            n_threads=0)
    assert np.array_equal(got.reversed, ref.reversed)
    for r in (got, ref):
        r.smoothed[np.isnan(r.smoothed)] = 0.0
        r.smoothed_cov[np.isnan(r.smoothed_cov)] = 0.0
    rep = error_report(got, ref, ev, max_chi2_ndf=100.0)
    print(rep, flush=True)
    assert rep["status_mismatch"] <= 0.005 * ev.n_tracks
    assert max(rep["smoothed_rms_rel"]) < 1e-6
    assert max(rep["smoothed_p999_rel"]) < 1e-5
    assert rep["smoothed_cov_max_rel"] < 1e-3
