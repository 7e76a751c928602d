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
