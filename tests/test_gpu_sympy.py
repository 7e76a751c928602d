"""GPU parity of the SympyStepper (SympyStepper.cpp:141-305): the HIP
fitter with stepper = SYMPY against the CPU oracle's SympyStepper on
identical inputs, at the FP64 bar (identical status and state sequence;
parameters and predicted/filtered covariances within 1e-10; smoothed
covariances within 5e-8 of sqrt(C_ii C_jj)), with either smoother; and the
transport skip leaves the results bit for bit unchanged.

Jacobians within 1e-8: the q/p column of the stepper's bound-to-free
jacobian takes the stages' field terms as dir3 - dir and dir4 - dir (the
form generate_sympy_stepper.py:364,372 writes), differences that cancel
down to the bend h q/p B / 2 and so magnify last-bit differences of the
trajectory (the two math libraries' sin/cos at the surfaces, ~5e-12) by
its inverse.  Measured in the field map: 2.3e-9 on d loc1 / d(q/p) and
d t / d(q/p) against 1.1e-11 with the EigenStepper, with the covariances
still within 3e-11.
"""
import numpy as np
import pytest

import acts_amd as A
from acts_amd import capi
from acts_amd.fitter import ALL_STATE_OUTPUTS

from parity import compare, cov_rel_err, error_report, param_rel_err, pulls

pytestmark = pytest.mark.gpu


def _fit(det, field, ev, opts):
    fitter = A.KalmanFitter(det, field, opts)
    try:
        return fitter.fit(ev, outputs=ALL_STATE_OUTPUTS)
    finally:
        fitter.close()


@pytest.mark.parametrize("workload,n,seed,smoother", [
    ("barrel", 1000, 111, capi.SMOOTHER_GAIN_MATRIX), ("endcap", 4000, 112, capi.SMOOTHER_MBF),
    ("solenoid", 3000, 113, capi.SMOOTHER_GAIN_MATRIX), ("ragged", 3000, 114, capi.SMOOTHER_MBF)])
def test_sympy_parity(oracle_mod, workload, n, seed, smoother):
    det, field, ev = A.workload(workload, n_tracks=n, seed=seed)
    opts = A.workload_options(workload, stepper=capi.STEPPER_SYMPY, smoother=smoother)
    print("sympy", workload, n, seed, smoother, flush=True)
    got = _fit(det, field, ev, opts)
    ref, _ = oracle_mod.fit(det, field, opts, ev, n_threads=0, outputs=ALL_STATE_OUTPUTS)
    rep = compare(got, ref, ev, 1e-10, rel_jacobian=1e-8)
    print(rep["max_rel_err"], flush=True)
    assert rep["ok"], rep
    assert np.count_nonzero(got.status == 0) >= 0.98 * n


def test_examples_configuration(oracle_mod):
    """The Examples' fit: SympyStepper, MbfSmoother and the fitted
    parameters extrapolated from the first measurement to a perigee at the
    origin, against the oracle in the same configuration."""
    det, field, ev = A.workload("solenoid", n_tracks=3000, seed=116)
    opts = A.KalmanFitterOptions(stepper=capi.STEPPER_SYMPY, smoother=capi.SMOOTHER_MBF,
                                 reference_surface=capi.REFERENCE_PERIGEE,
                                 reference_strategy=capi.EXTRAPOLATE_FIRST)
    out = ALL_STATE_OUTPUTS + ("reference_params", "reference_cov")
    fitter = A.KalmanFitter(det, field, opts)
    try:
        got = fitter.fit(ev, outputs=out)
    finally:
        fitter.close()
    ref, _ = oracle_mod.fit(det, field, opts, ev, n_threads=0, outputs=out)
    rep = compare(got, ref, ev, 1e-10, rel_jacobian=1e-8)
    print(rep["max_rel_err"], flush=True)
    assert rep["ok"], rep
    ok = (got.status == 0) & (ref.status == 0)
    assert param_rel_err(got.reference_params[ok], ref.reference_params[ok]).max() <= 1e-10
    assert cov_rel_err(got.reference_cov[ok], ref.reference_cov[ok]).max() <= 5e-8


def test_sympy_mixed_precision(oracle_mod):
    """The SympyStepper with FP32 stages and jacobian recursion against the
    FP64 oracle's SympyStepper, to the mixed-precision bounds of the
    EigenStepper (tests/test_gpu_mixed.py): status of >= 99.5 % of the tracks,
    smoothed-parameter RMS < 1e-6 and 99.9th percentile < 1e-5 over the
    converged fits, covariances within 1e-3, pulls as the FP64 fit's."""
    det, field, ev = A.workload("solenoid", n_tracks=6000, seed=117)
    got = _fit(det, field, ev, A.KalmanFitterOptions(stepper=capi.STEPPER_SYMPY,
                                                     precision=capi.PRECISION_MIXED))
    ref, _ = oracle_mod.fit(det, field, A.KalmanFitterOptions(stepper=capi.STEPPER_SYMPY), ev,
                            n_threads=0)
    rep = error_report(got, ref, ev, max_chi2_ndf=100.0)
    pg, pr = pulls(got, ev), pulls(ref, ev)
    print(rep, flush=True)
    assert rep["status_mismatch"] <= 0.005 * ev.n_tracks
    assert max(rep["smoothed_rms_rel"]) < 1e-6
    assert max(rep["smoothed_p999_rel"]) < 1e-5
    assert rep["smoothed_cov_max_rel"] < 1e-3
    assert np.all(np.abs(np.array(pg["mean"]) - pr["mean"]) < 0.002)
    assert np.all(np.abs(np.array(pg["std"]) / pr["std"] - 1) < 0.002)


def test_sympy_transport_skip_bitwise():
    det, field, ev = A.workload("solenoid", n_tracks=2000, seed=115)
    out = {}
    for skip in (0, 1, 64):
        out[skip] = _fit(det, field, ev, A.KalmanFitterOptions(stepper=capi.STEPPER_SYMPY,
                                                               sched_transport_skip_steps=skip))
    for skip in (1, 64):
        for name in ("status", "n_steps", "smoothed", "smoothed_cov", "predicted_cov", "jacobian"):
            assert np.array_equal(getattr(out[skip], name), getattr(out[0], name)), (skip, name)
