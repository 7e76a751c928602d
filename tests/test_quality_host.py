"""The batch fit-quality figures of acts_amd.quality (used by the bench):
the vectorised first-state pulls agree with the per-track pull loop of the
tests, and the relative error finds planted differences and sets
non-converged fits apart."""
import numpy as np

import acts_amd as A
from acts_amd import quality
from acts_amd.fitter import ALL_STATE_OUTPUTS

from parity import pulls as loop_pulls


class _HostResult:
    """A FitResult seen through the DeviceResult accessors."""

    def __init__(self, r):
        self.t = {k: getattr(r, k) for k in ("status", "n_states", "chi2", "ndf",
                                             "state_surface", "state_flags", "smoothed",
                                             "smoothed_cov")}

    def host(self, name):
        return np.asarray(self.t[name])

    def host_rows(self, name, rows):
        return np.asarray(self.t[name])[rows]


def _fit(oracle_mod, seed):
    det, field, ev = A.workload("barrel", n_tracks=1500, seed=seed)
    r, _ = oracle_mod.fit(det, field, A.KalmanFitterOptions(), ev, n_threads=0,
                          outputs=ALL_STATE_OUTPUTS)
    return ev, r


def test_vectorised_pulls_match_loop(oracle_mod):
    ev, r = _fit(oracle_mod, 71)
    q = quality.pulls(quality.first_states(_HostResult(r), ev), ev)
    p = loop_pulls(r, ev)
    assert q["n"] == p["n"] and q["n_beyond"] == p["n_beyond_10"]
    np.testing.assert_allclose(list(q["mean"].values()), p["mean"], atol=1e-4)
    np.testing.assert_allclose(list(q["std"].values()), p["std"], atol=1e-4)


def test_relative_error_finds_planted_difference(oracle_mod):
    ev, r = _fit(oracle_mod, 72)
    ref = quality.first_states(_HostResult(r), ev)
    same = quality.relative_error(ref, ref)
    assert same["max"] == 0.0 and same["status_mismatch"] == 0
    got = {k: v.copy() for k, v in ref.items()}
    t = int(np.nonzero(ref["status"] == 0)[0][3])
    got["smoothed"][t, 4] *= 1 + 1e-6          # q/p, relative 1e-6
    rep = quality.relative_error(got, ref)
    assert abs(rep["max"] - 1e-6) < 1e-9
    # a fit with chi2/ndf above the cut is counted apart, not compared
    ref2 = {k: v.copy() for k, v in ref.items()}
    ref2["chi2"] = ref2["chi2"].copy()
    ref2["chi2"][t] = 1e9
    rep2 = quality.relative_error(got, ref2)
    assert rep2["max"] == 0.0 and rep2["n_not_converged"] == 1
    assert abs(rep2["max_incl_not_converged"] - 1e-6) < 1e-9
