"""Comparison of two batched fit results (HIP fitter vs CPU oracle).

Parameters are compared with |a - b| <= rel * max(|b|, floor) per bound
component (floors keep near-zero values from demanding absolute agreement
below the parameters' natural scale); covariances with
|a_ij - b_ij| <= rel * sqrt(|b_ii b_jj|), except the filtered covariance,
which is measured against sqrt(|p_ii p_jj|) of the predicted covariance
p it is computed from: the update C_f = C_p - K H C_p cancels down to
the measurement's resolution, and an update that shrinks a variance by
~1e5 (a precise hit after a long unmeasured stretch at high momentum)
leaves last-bit differences of C_p amplified that much relative to C_f
(measured: 7.2e-9 relative to C_f, 3e-14 relative to C_p, on the last
state of one 86 GeV track in 3000 in the field map).

Tolerances of the FP64 fit: 1e-10 for every parameter vector, the
predicted/filtered covariances and the transport jacobians; 5e-8 for the
smoothed covariance, whose RTS update inverts the predicted covariance
(condition numbers up to ~1e14 in the mixed mm/rad/GeV^-1/ns units at
100 GeV) and so amplifies last-bit differences of the double-precision
sin/cos/atan2 of the two math libraries by up to ~1e4 (measured: 3.4e-10
at 1-10 GeV; 5.4e-9 and 1.5e-8 for two samples of 4000 tracks at
0.5-100 GeV in the field map, with predicted and filtered covariances
within 3e-12).
"""
from __future__ import annotations

import numpy as np

# loc0 [mm], loc1 [mm], phi, theta [rad], q/p [1/GeV], time [mm]
PARAM_FLOORS = np.array([1.0, 1.0, 1.0, 1.0, 0.01, 1.0])


def param_rel_err(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    scale = np.maximum(np.abs(b), PARAM_FLOORS)
    return np.abs(a - b) / scale


def cov_rel_err(a: np.ndarray, b: np.ndarray, scale_cov: np.ndarray | None = None) -> np.ndarray:
    """|a_ij - b_ij| / sqrt(|s_ii s_jj|) with s = scale_cov (default b)."""
    a = a.reshape(-1, 6, 6)
    b = b.reshape(-1, 6, 6)
    s = b if scale_cov is None else scale_cov.reshape(-1, 6, 6)
    d = np.sqrt(np.abs(np.einsum("nii->ni", s)))
    scale = d[:, :, None] * d[:, None, :]
    scale = np.where(scale > 0, scale, 1.0)
    return (np.abs(a - b) / scale).reshape(len(a), -1)


def compare(got, ref, events, rel: float = 1e-10, rel_smoothed_cov: float = 5e-8,
            rel_jacobian: float | None = None, rel_cov: float | None = None,
            max_chi2_ndf: float | None = None) -> dict:
    """Return a report; `ok` is True if everything agrees within tolerance.
    With `max_chi2_ndf`, tracks whose reference fit has a larger chi2/ndf
    (fits that did not converge, e.g. after a catastrophic energy loss on
    the last surfaces, whose smoother amplifies last-bit differences far
    beyond the bar) must still agree in status, states and counters, but
    are left out of the parameter and covariance bars (`n_not_converged`)."""
    n = events.n_tracks
    rep = {"n_tracks": n}
    same_status = got.status == ref.status
    rep["status_mismatch"] = int(np.count_nonzero(~same_status))
    both_ok = same_status & (ref.status == 0)
    rep["n_ok"] = int(np.count_nonzero(both_ok))
    rep["n_failed_ref"] = int(np.count_nonzero(ref.status != 0))
    same_ns = got.n_states == ref.n_states
    rep["n_states_mismatch"] = int(np.count_nonzero(both_ok & ~same_ns))
    good = both_ok & same_ns
    rep["n_not_converged"] = 0
    for name in ("chi2", "ndf", "n_measurements", "n_holes", "n_outliers"):
        g, r = getattr(got, name), getattr(ref, name)
        if name == "chi2":
            bad = good & (np.abs(g - r) > 1e-5 * np.maximum(np.abs(r), 1.0))
        else:
            bad = good & (g != r)
        rep[f"{name}_mismatch"] = int(np.count_nonzero(bad))
    # per-state arrays over the states of the good tracks
    so = events.seq_offsets
    idx = np.concatenate([np.arange(so[t], so[t] + ref.n_states[t])
                          for t in np.nonzero(good)[0]]) if good.any() else np.zeros(0, int)
    seq_bad_slots = np.zeros(0, bool)
    if ref.state_surface is not None and got.state_surface is not None:
        seq_bad_slots = (got.state_surface[idx] != ref.state_surface[idx])
        if ref.state_flags is not None and got.state_flags is not None:
            seq_bad_slots |= got.state_flags[idx] != ref.state_flags[idx]
    rep["state_sequence_mismatch_slots"] = int(np.count_nonzero(seq_bad_slots))
    if max_chi2_ndf is not None:
        converged = ref.chi2 / np.maximum(ref.ndf, 1) <= max_chi2_ndf
        rep["n_not_converged"] = int(np.count_nonzero(good & ~converged))
        good = good & converged
        idx = np.concatenate([np.arange(so[t], so[t] + ref.n_states[t])
                              for t in np.nonzero(good)[0]]) if good.any() else np.zeros(0, int)
    worst = {}
    for name in ("predicted", "filtered", "smoothed"):
        g, r = getattr(got, name), getattr(ref, name)
        if g is not None and r is not None and len(idx):
            worst[name] = float(np.max(param_rel_err(g[idx], r[idx])))
    for name in ("predicted_cov", "filtered_cov", "smoothed_cov", "jacobian"):
        g, r = getattr(got, name), getattr(ref, name)
        if g is not None and r is not None and len(idx):
            if name == "jacobian":
                worst[name] = float(np.max(np.abs(g[idx] - r[idx]) /
                                           np.maximum(np.abs(r[idx]), 1.0)))
            elif name == "filtered_cov" and ref.predicted_cov is not None:
                # the update C_f = C_p - K H C_p subtracts operands of the
                # size of C_p: its rounding is relative to C_p
                worst[name] = float(np.max(cov_rel_err(g[idx], r[idx], ref.predicted_cov[idx])))
            else:
                worst[name] = float(np.max(cov_rel_err(g[idx], r[idx])))
    rep["max_rel_err"] = worst
    # tracks whose smoothed parameters and covariances are bit-identical
    if ref.smoothed is not None and got.smoothed is not None and len(idx):
        slot_exact = np.all(got.smoothed[idx] == ref.smoothed[idx], axis=1)
        if ref.smoothed_cov is not None and got.smoothed_cov is not None:
            slot_exact &= np.all(got.smoothed_cov[idx] == ref.smoothed_cov[idx], axis=1)
        owner = np.repeat(np.nonzero(good)[0], ref.n_states[good])
        bad_tracks = np.unique(owner[~slot_exact])
        rep["n_bitwise_identical_tracks"] = int(np.count_nonzero(good)) - len(bad_tracks)
    if ref.state_chi2 is not None and got.state_chi2 is not None and len(idx):
        rep["max_state_chi2_rel"] = float(np.max(
            np.abs(got.state_chi2[idx].astype(np.float64) - ref.state_chi2[idx]) /
            np.maximum(np.abs(ref.state_chi2[idx]), 1.0)))
    rep["ok"] = (rep["status_mismatch"] == 0 and rep["n_states_mismatch"] == 0
                 and rep["state_sequence_mismatch_slots"] == 0
                 and all(v <= (max(rel_smoothed_cov, rel_cov or 0) if k == "smoothed_cov" else
                               rel_jacobian if k == "jacobian" and rel_jacobian else
                               rel_cov if k.endswith("_cov") and rel_cov else rel)
                         for k, v in worst.items())
                 and all(rep[f"{k}_mismatch"] == 0 for k in
                         ("chi2", "ndf", "n_measurements", "n_holes", "n_outliers")))
    return rep


def float_qop_flips(got, ref, events) -> np.ndarray:
    """Tracks (both fitted, same state count) with a state whose filtered
    q/p rounds to different floats in the two fits.  The material effects
    are computed in float from the q/p, so at such a state the process
    noise of the next surface moves by a float ulp and every covariance
    after it by ~1e-8, while the double-precision parameters still agree
    to ~1e-13."""
    so = events.seq_offsets
    flips = np.zeros(events.n_tracks, bool)
    for t in np.nonzero((got.status == 0) & (ref.status == 0) & (got.n_states == ref.n_states))[0]:
        s = slice(so[t], so[t] + ref.n_states[t])
        flips[t] = np.any(got.filtered[s, 4].astype(np.float32) != ref.filtered[s, 4].astype(np.float32))
    return flips


def smoothing_spread(rts, mbf, events) -> np.ndarray:
    """Per track, the largest relative difference (floors as in `compare`)
    between the smoothed parameters of the oracle's RTS and MBF smoothers on
    the same forward pass.  The two recursions are equal in exact arithmetic,
    so the difference measures how well the track's smoothed parameters are
    determined in double precision at all: ~1e-12 for a converged fit, O(1)
    for a fit that did not converge (chi2/ndf of 1e6 and more after a
    residual across the +-pi wrap of loc0 = R phi), whose smoothing inverts
    predicted covariances the residuals disagree with by orders of
    magnitude.  Tracks that either fit failed get 0."""
    so = events.seq_offsets
    spread = np.zeros(events.n_tracks)
    both = (rts.status == 0) & (mbf.status == 0) & (rts.n_states == mbf.n_states)
    for t in np.nonzero(both)[0]:
        s = slice(so[t], so[t] + rts.n_states[t])
        spread[t] = float(param_rel_err(mbf.smoothed[s], rts.smoothed[s]).max())
    return spread


def error_report(got, ref, events, max_chi2_ndf: float | None = None) -> dict:
    """Errors of a reduced-precision fit against the FP64 reference fit.

    Over the tracks that both fit with the same state sequence: max and RMS
    of the relative parameter error (floors as in `compare`) of the smoothed
    parameters, and of the smoothed covariance (relative to sqrt(C_ii C_jj)).
    With `max_chi2_ndf`, only tracks whose reference fit has chi2/ndf at most
    that are compared (a fit that did not converge, e.g. after a residual
    across the +-pi wrap of loc0 = R phi, has no meaningful parameters).
    """
    same = (got.status == ref.status) & (ref.status == 0) & (got.n_states == ref.n_states)
    n_same = int(np.count_nonzero(same))
    if max_chi2_ndf is not None:
        same &= ref.chi2 / np.maximum(ref.ndf, 1) <= max_chi2_ndf
    so = events.seq_offsets
    idx = np.concatenate([np.arange(so[t], so[t] + ref.n_states[t])
                          for t in np.nonzero(same)[0]]) if same.any() else np.zeros(0, int)
    if len(idx) and got.state_surface is not None:
        seq_ok = np.ones(len(idx), bool)
        seq_ok &= got.state_surface[idx] == ref.state_surface[idx]
        idx = idx[seq_ok]
    rep = {"n_tracks": int(events.n_tracks),
           "status_mismatch": int(np.count_nonzero(got.status != ref.status)),
           "n_compared_states": int(len(idx)),
           "n_not_converged": n_same - int(np.count_nonzero(same))}
    if len(idx):
        pe = param_rel_err(got.smoothed[idx], ref.smoothed[idx])
        rep["smoothed_max_rel"] = pe.max(axis=0).tolist()
        rep["smoothed_p999_rel"] = np.quantile(pe, 0.999, axis=0).tolist()
        rep["smoothed_rms_rel"] = np.sqrt((pe ** 2).mean(axis=0)).tolist()
        if got.smoothed_cov is not None:
            ce = cov_rel_err(got.smoothed_cov[idx], ref.smoothed_cov[idx])
            rep["smoothed_cov_max_rel"] = float(ce.max())
            rep["smoothed_cov_rms_rel"] = float(np.sqrt((ce ** 2).mean()))
    return rep


def pulls(res, events) -> dict:
    """Pull mean and width of the smoothed parameters (loc0, loc1, phi,
    theta, q/p) at each track's first measurement state, against the
    generator's truth there."""
    so, mo = events.seq_offsets, events.meas_offsets
    vals = []
    for t in np.nonzero(res.status == 0)[0]:
        for k in range(res.n_states[t]):
            s = so[t] + k
            if res.state_flags[s] & 1:  # HAS_MEASUREMENT
                surf = res.state_surface[s]
                ms = np.nonzero(events.meas_surfaces[mo[t]:mo[t + 1]] == surf)[0]
                if not len(ms):
                    break
                truth = events.truth_meas[mo[t] + ms[0]]
                cov = res.smoothed_cov[s].reshape(6, 6)
                d = res.smoothed[s] - truth
                d[2] = (d[2] + np.pi) % (2 * np.pi) - np.pi
                sig = np.sqrt(np.diag(cov))
                vals.append((d / np.where(sig > 0, sig, 1.0))[:5])
                break
    v = np.array(vals)
    # a track whose fitted position sits across the +-pi seam of a cylinder
    # (loc0 = R phi) has a pull of ~2 pi R / sigma: keep |pull| < 10 and
    # report how many are beyond
    keep = np.all(np.abs(v) < 10, axis=1)
    w = v[keep]
    return {"n": int(len(v)), "n_beyond_10": int(np.count_nonzero(~keep)),
            "mean": w.mean(axis=0).tolist(), "std": w.std(axis=0).tolist()}
