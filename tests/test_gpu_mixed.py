"""Mixed precision (FP32 stepping and FP32 field map, FP64 Kalman algebra)
against the FP64 oracle on the same inputs.

Stated accuracy (checked here): the same fit status for >= 99.5 % of the
tracks; over the states of the tracks with the same state sequence whose
FP64 fit converged (chi2/ndf <= 100), the RMS of the relative error of the
smoothed parameters (floors 1 mm, 1 mm, 1 rad, 1 rad, 0.01/GeV, 1 mm) below
1e-6 and its 99.9th percentile below 1e-5, smoothed covariances within
1e-3 of sqrt(C_ii C_jj); pull means within 0.002 and pull widths within
0.2 % of the FP64 fit's.  The maximum error is reported (measured 4.5e-6 on
q/p and 5e-5 on the unmeasured time over 77 000 states of 6000 tracks;
RMS <= 2e-7, 99.9th percentile <= 1e-6).
"""
import numpy as np
import pytest

import acts_amd as A
from acts_amd import capi
from acts_amd.fitter import ALL_STATE_OUTPUTS

from parity import error_report, pulls

pytestmark = pytest.mark.gpu


def test_mixed_precision_solenoid(oracle_mod):
    det, field, ev = A.workload("solenoid", n_tracks=6000, seed=21)
    fitter = A.KalmanFitter(det, field, A.KalmanFitterOptions(precision=capi.PRECISION_MIXED))
    try:
        got = fitter.fit(ev, outputs=ALL_STATE_OUTPUTS)
    finally:
        fitter.close()
    ref, _ = oracle_mod.fit(det, field, A.KalmanFitterOptions(), ev, n_threads=0)
    rep = error_report(got, ref, ev, max_chi2_ndf=100.0)
    pg, pr = pulls(got, ev), pulls(ref, ev)
    print(rep, flush=True)
    print("pulls mixed", pg, flush=True)
    print("pulls fp64 ", pr, flush=True)
    assert rep["status_mismatch"] <= 0.005 * ev.n_tracks
    assert max(rep["smoothed_rms_rel"]) < 1e-6
    assert max(rep["smoothed_p999_rel"]) < 1e-5
    assert rep["smoothed_cov_max_rel"] < 1e-3
    assert np.all(np.abs(np.array(pg["mean"]) - pr["mean"]) < 0.002)
    assert np.all(np.abs(np.array(pg["std"]) / pr["std"] - 1) < 0.002)
