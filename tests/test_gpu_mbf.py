"""GPU parity of the MbfSmoother (MbfSmoother.cpp:17-121): the HIP fitter
with KalmanFitterOptions smoother = MBF against the CPU oracle's MBF on
identical inputs, to the FP64 bar (identical status and state sequence;
parameters, predicted/filtered covariances and jacobians within 1e-10;
smoothed covariances within 5e-8 of sqrt(C_ii C_jj)); the perigee
extrapolation from the MBF-smoothed first state to the same bar; the
mixed-precision fit with MBF within the mixed-precision bounds.
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


@pytest.mark.parametrize("workload,n,seed", [("barrel", 1000, 41), ("endcap", 4000, 42),
                                             ("solenoid", 3000, 43), ("ragged", 4000, 44)])
def test_mbf_parity(oracle_mod, workload, n, seed):
    det, field, ev = A.workload(workload, n_tracks=n, seed=seed)
    kw = dict(WORKLOADS[workload].get("options", {}))
    kw.update(smoother=capi.SMOOTHER_MBF, reference_surface=capi.REFERENCE_PERIGEE,
              reference_strategy=capi.EXTRAPOLATE_FIRST)
    print("mbf", workload, n, seed, flush=True)
    got = _fit(det, field, ev, **kw)
    ref, _ = oracle_mod.fit(det, field, A.KalmanFitterOptions(**kw), ev, n_threads=0,
                            outputs=OUTPUTS)
    rep = compare(got, ref, ev, 1e-10)
    print(rep["max_rel_err"], flush=True)
    assert rep["ok"], rep
    ok = (got.status == 0) & (ref.status == 0)
    assert np.count_nonzero(ok) >= 0.98 * n
    assert param_rel_err(got.reference_params[ok], ref.reference_params[ok]).max() <= 1e-10
    assert cov_rel_err(got.reference_cov[ok], ref.reference_cov[ok]).max() <= 5e-8


def test_mbf_mixed_precision(oracle_mod):
    det, field, ev = A.workload("solenoid", n_tracks=4000, seed=45)
    got = _fit(det, field, ev, smoother=capi.SMOOTHER_MBF, precision=capi.PRECISION_MIXED)
    ref, _ = oracle_mod.fit(det, field, A.KalmanFitterOptions(smoother=capi.SMOOTHER_MBF), ev,
                            n_threads=0)
    rep = error_report(got, ref, ev, max_chi2_ndf=100.0)
    print(rep, flush=True)
    assert rep["status_mismatch"] <= 0.005 * ev.n_tracks
    assert max(rep["smoothed_rms_rel"]) < 1e-6
    assert max(rep["smoothed_p999_rel"]) < 1e-5
    assert rep["smoothed_cov_max_rel"] < 1e-3

