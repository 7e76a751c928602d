"""Fit-quality figures of a batch: pulls against the generator's truth and
the error of one fit of a batch against another (e.g. mixed precision
against FP64).

Both work on the first state of every track only, gathered on the device,
so that they stay cheap for batches of millions of tracks.
"""
from __future__ import annotations

import numpy as np

from . import capi

# loc0 [mm], loc1 [mm], phi, theta [rad], q/p [1/GeV], time [mm]
PARAM_FLOORS = np.array([1.0, 1.0, 1.0, 1.0, 0.01, 1.0])
PULL_NAMES = ("loc0", "loc1", "phi", "theta", "qop")


def first_states(res, events, names=("state_surface", "state_flags", "smoothed",
                                      "smoothed_cov")) -> dict:
    """The first-state rows of the per-state outputs `names` of a
    DeviceResult, plus the per-track summaries."""
    rows = events.seq_offsets[:-1]
    out = {name: res.host_rows(name, rows) for name in names if name in res.t}
    for name in ("status", "n_states", "chi2", "ndf"):
        out[name] = res.host(name)
    return out


def pulls(first: dict, events, limit: float = 10.0) -> dict:
    """Pull mean and width of the smoothed (loc0, loc1, phi, theta, q/p) at
    each track's first state, where that state holds the track's first
    measurement, against the generator's truth there.  Fits whose pull
    exceeds `limit` in some component (e.g. a position fitted across the
    +-pi seam of a cylinder, ~2 pi R / sigma) are counted, not averaged."""
    mo = events.meas_offsets
    n = events.n_tracks
    has_meas = (first["state_flags"] & capi.STATE_HAS_MEASUREMENT) != 0
    nm = np.diff(mo) > 0
    m0 = np.where(nm, mo[:-1], 0)
    ok = (first["status"] == 0) & has_meas & nm
    ok &= events.meas_surfaces[m0] == first["state_surface"]
    t = np.nonzero(ok)[0]
    truth = events.truth_meas[m0[t]]
    d = first["smoothed"][t] - truth
    d[:, 2] = (d[:, 2] + np.pi) % (2 * np.pi) - np.pi
    var = first["smoothed_cov"][t].reshape(-1, 6, 6)[:, np.arange(6), np.arange(6)]
    sig = np.sqrt(np.abs(var))
    p = (d / np.where(sig > 0, sig, 1.0))[:, :5]
    keep = np.all(np.abs(p) < limit, axis=1)
    w = p[keep]
    return {"n": int(len(p)), "n_beyond": int(np.count_nonzero(~keep)), "limit": limit,
            "of": n,
            "mean": dict(zip(PULL_NAMES, np.round(w.mean(axis=0), 4).tolist())),
            "std": dict(zip(PULL_NAMES, np.round(w.std(axis=0), 4).tolist()))}


def relative_error(got: dict, ref: dict, max_chi2_ndf: float = 100.0) -> dict:
    """Relative error |got - ref| / max(|ref|, floor) of the first state's
    smoothed parameters over the tracks that both fit (status 0, same number
    of states, same first surface) and whose reference fit converged
    (chi2/ndf <= max_chi2_ndf).  A fit that did not converge (chi2/ndf of
    1e6 and more: a residual across a +-pi wrap of loc0 = R phi or of a
    disc's phi) has no meaningful parameters, and its smoothed values change
    by O(1) with the last bit of any input; those are counted apart."""
    both = (got["status"] == 0) & (ref["status"] == 0) & (got["n_states"] == ref["n_states"])
    both &= got["state_surface"] == ref["state_surface"]
    chi2n = ref["chi2"] / np.maximum(ref["ndf"], 1)
    conv = both & (chi2n <= max_chi2_ndf)
    e = np.abs(got["smoothed"] - ref["smoothed"]) / np.maximum(np.abs(ref["smoothed"]),
                                                                PARAM_FLOORS)
    ec = e[conv].max(axis=1)
    ea = e[both].max(axis=1)
    return {"n_compared": int(np.count_nonzero(conv)),
            "n_not_converged": int(np.count_nonzero(both & ~conv)),
            "max_chi2_ndf": max_chi2_ndf,
            "status_mismatch": int(np.count_nonzero(got["status"] != ref["status"])),
            "rms": float(np.sqrt((ec ** 2).mean())) if len(ec) else None,
            "p999": float(np.quantile(ec, 0.999)) if len(ec) else None,
            "max": float(ec.max()) if len(ec) else None,
            "max_incl_not_converged": float(ea.max()) if len(ea) else None}


def parity_counts(got, ref, seq_offsets, rel: float = 1e-10, rel_cov: float = 5e-8,
                  chunk: int = 1 << 20) -> dict:
    """Agreement of two fits of one batch (objects with the per-track
    summaries and the per-state arrays state_surface, state_flags, smoothed
    and optionally smoothed_cov), e.g. the GPU fit against the CPU fit:
    tracks with another status or state count, slots of the compared tracks
    with another (surface, flags), and the largest relative errors of the
    smoothed parameters (|a - b| <= rel * max(|b|, floor)) and covariances
    (|a_ij - b_ij| <= rel_cov * sqrt(b_ii b_jj)) with the number of tracks
    beyond each bar.  Vectorised over slots in chunks, for whole batches."""
    n = len(ref.status)
    same = (got.status == ref.status) & (got.n_states == ref.n_states)
    good = same & (ref.status == 0)
    so = np.asarray(seq_offsets)
    lens = np.where(good, ref.n_states, 0).astype(np.int64)
    owner = np.repeat(np.arange(n), lens)
    starts = np.repeat(so[:-1], lens)
    within = np.arange(len(owner)) - np.repeat(np.cumsum(lens) - lens, lens)
    idx = starts + within
    out = {"n_tracks": int(n), "n_compared_tracks": int(np.count_nonzero(good)),
           "n_compared_states": int(len(idx)),
           "status_or_state_count_mismatch": int(np.count_nonzero(~same)),
           "rel_bar": rel, "rel_cov_bar": rel_cov}
    seq_bad = np.zeros(n, bool)
    p_bad = np.zeros(n, bool)
    c_bad = np.zeros(n, bool)
    p_max = 0.0
    c_max = 0.0
    has_cov = getattr(got, "smoothed_cov", None) is not None and \
        getattr(ref, "smoothed_cov", None) is not None
    for a in range(0, len(idx), chunk):
        ii, oo = idx[a:a + chunk], owner[a:a + chunk]
        bad = (got.state_surface[ii] != ref.state_surface[ii]) | \
              (got.state_flags[ii] != ref.state_flags[ii])
        seq_bad[oo[bad]] = True
        pe = (np.abs(got.smoothed[ii] - ref.smoothed[ii]) /
              np.maximum(np.abs(ref.smoothed[ii]), PARAM_FLOORS)).max(axis=1)
        p_max = max(p_max, float(pe.max()) if len(pe) else 0.0)
        p_bad[oo[pe > rel]] = True
        if has_cov:
            g = got.smoothed_cov[ii].reshape(-1, 6, 6)
            r = ref.smoothed_cov[ii].reshape(-1, 6, 6)
            d = np.sqrt(np.abs(np.einsum("nii->ni", r)))
            sc = d[:, :, None] * d[:, None, :]
            ce = (np.abs(g - r) / np.where(sc > 0, sc, 1.0)).reshape(len(ii), -1).max(axis=1)
            c_max = max(c_max, float(ce.max()) if len(ce) else 0.0)
            c_bad[oo[ce > rel_cov]] = True
    out["state_sequence_mismatch_tracks"] = int(np.count_nonzero(seq_bad))
    out["smoothed_max_rel"] = p_max
    out["tracks_beyond_rel_bar"] = int(np.count_nonzero(p_bad))
    if has_cov:
        out["smoothed_cov_max_rel"] = c_max
        out["tracks_beyond_rel_cov_bar"] = int(np.count_nonzero(c_bad))
    chi2_bad = good & (np.abs(got.chi2 - ref.chi2) > 1e-5 * np.maximum(np.abs(ref.chi2), 1.0))
    out["chi2_mismatch_tracks"] = int(np.count_nonzero(chi2_bad))
    return out
