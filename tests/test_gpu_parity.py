"""GPU parity: the HIP fitter against the CPU oracle on identical inputs.

Bar: identical per-track status, state count, surface/type sequence and
track summary; parameters and covariances within 1e-10 relative
(tests/parity.py); chi2 (float) within float rounding.
"""
import numpy as np
import pytest

import acts_amd as A
from acts_amd.fitter import ALL_STATE_OUTPUTS

from parity import compare

pytestmark = pytest.mark.gpu

REL = 1e-10


def _fit_both(oracle_mod, det, field, ev, opts=None):
    opts = opts or A.KalmanFitterOptions()
    fitter = A.KalmanFitter(det, field, opts)
    try:
        got = fitter.fit(ev, outputs=ALL_STATE_OUTPUTS)
    finally:
        fitter.close()
    ref, _ = oracle_mod.fit(det, field, opts, ev, n_threads=0)
    return got, ref


def _report(rep):
    print({k: v for k, v in rep.items()}, flush=True)


def test_barrel_1000_tracks(oracle_mod):
    """1,000 muons, 10 barrel layers, Bz = 2 T, |eta| < 1."""
    det, field, ev = A.workload("barrel", n_tracks=1000, seed=11)
    got, ref = _fit_both(oracle_mod, det, field, ev)
    rep = compare(got, ref, ev, REL)
    _report(rep)
    assert rep["n_ok"] >= 990
    assert rep["ok"], rep


def test_endcap_tracks(oracle_mod):
    """Barrel + 2x7 discs, |eta| < 2.5: 6-17 measurements per track."""
    det, field, ev = A.workload("endcap", n_tracks=4000, seed=12)
    got, ref = _fit_both(oracle_mod, det, field, ev)
    rep = compare(got, ref, ev, REL)
    _report(rep)
    assert rep["ok"], rep


def test_solenoid_field_map(oracle_mod):
    """Barrel + discs in a 201x201x401 map of a 2 T solenoid (trilinear
    interpolation), pT 0.5-100 GeV; includes tracks that leave the map."""
    det, field, ev = A.workload("solenoid", n_tracks=4000, seed=13)
    got, ref = _fit_both(oracle_mod, det, field, ev)
    rep = compare(got, ref, ev, REL)
    _report(rep)
    assert rep["n_ok"] >= 3900
    assert rep["ok"], rep
