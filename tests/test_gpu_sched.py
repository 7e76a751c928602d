"""Scheduling knobs change only the speed of a fit: results are bit for bit
the same with quarantine, other inner step counts and wait fractions, and in
input order."""
import numpy as np
import pytest

import acts_amd as A
from acts_amd.fitter import ALL_STATE_OUTPUTS

pytestmark = pytest.mark.gpu

VARIANTS = [
    dict(sched_quarantine_steps=100),
    dict(sched_quarantine_steps=20, sched_inner_steps=64, sched_wait_64ths=64),
    dict(sched_inner_steps=4, sched_wait_64ths=16),
    dict(length_sorted=False),
]


def test_scheduling_does_not_change_results():
    det, field, ev = A.workload("endcap", n_tracks=5000, seed=51)
    f0 = A.KalmanFitter(det, field, A.KalmanFitterOptions())
    try:
        ref = f0.fit(ev, outputs=ALL_STATE_OUTPUTS)
    finally:
        f0.close()
    for kw in VARIANTS:
        print("variant", kw, flush=True)
        f = A.KalmanFitter(det, field, A.KalmanFitterOptions(**kw))
        try:
            got = f.fit(ev, outputs=ALL_STATE_OUTPUTS)
        finally:
            f.close()
        for name in ("status", "n_states", "chi2", "n_steps", "n_step_trials", "smoothed",
                     "smoothed_cov", "predicted", "filtered_cov", "jacobian", "state_path"):
            assert np.array_equal(getattr(got, name), getattr(ref, name)), (kw, name)
