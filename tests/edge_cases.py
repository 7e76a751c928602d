"""Edge-case batches for the fit tests, built from ordinary tracks: no
measurements, an empty surface sequence, a single measurement, a hole (a
measurement taken away, its surface kept), a measurement on a surface that
the track's sequence does not contain (never reached) together with the
same track without that measurement, and measurements only on the last
two surfaces, followed by ordinary tracks."""
from __future__ import annotations

import numpy as np

from acts_amd.events import Events

# (kind, index of the source track)
CASES = (("no measurements", 0), ("empty sequence", 1), ("single measurement", 2),
         ("hole", 3), ("unreachable measurement", 4), ("unreachable removed", 4),
         ("last two measurements only", 5))


def _edit(kind, seq, mi, ev):
    """The sequence and measurement indices of one edge case."""
    if kind == "no measurements":
        return seq, []
    if kind == "empty sequence":
        return [], []
    if kind == "single measurement":
        return seq, mi[:1]
    if kind == "hole":
        return seq, mi[:len(mi) // 2] + mi[len(mi) // 2 + 1:]
    if kind in ("unreachable measurement", "unreachable removed"):
        # the middle measurement's surface taken out of the sequence; the
        # first kind keeps the measurement in the input, the second drops it
        m = mi[len(mi) // 2]
        gone = ev.meas_surfaces[m]
        seq = [s for s in seq if s != gone]
        return seq, (mi if kind == "unreachable measurement" else [x for x in mi if x != m])
    if kind == "last two measurements only":
        return seq, mi[-2:]
    return seq, mi


def edge_batch(ev: Events, n_ordinary: int = 10) -> tuple[Events, tuple]:
    """A batch of the edge cases followed by `n_ordinary` unchanged tracks
    (sources len(CASES) .. len(CASES)+n_ordinary-1); returns it and the kind
    of each track."""
    so, mo = ev.seq_offsets, ev.meas_offsets
    cases = list(CASES) + [("ordinary", len(CASES) + k) for k in range(n_ordinary)]
    seqs, meas = [], []
    for kind, t in cases:
        seq, mi = _edit(kind, list(ev.seq_surfaces[so[t]:so[t + 1]]), list(range(mo[t], mo[t + 1])), ev)
        seqs.append(np.asarray(seq, np.int32))
        meas.append(np.asarray(mi, np.int64))
    src = np.asarray([t for _, t in cases], np.int64)
    n = len(cases)
    new_so = np.zeros(n + 1, np.int64)
    new_so[1:] = np.cumsum([len(s) for s in seqs])
    new_mo = np.zeros(n + 1, np.int64)
    new_mo[1:] = np.cumsum([len(m) for m in meas])
    mid = np.concatenate(meas)
    out = Events(
        n_tracks=n,
        start_params=np.ascontiguousarray(ev.start_params[src]),
        start_cov=np.ascontiguousarray(ev.start_cov[src]),
        start_plane=np.ascontiguousarray(ev.start_plane[src]),
        seq_offsets=new_so,
        seq_surfaces=np.ascontiguousarray(np.concatenate(seqs).astype(np.int32)),
        meas_offsets=new_mo,
        meas_surfaces=np.ascontiguousarray(ev.meas_surfaces[mid]),
        meas_values=np.ascontiguousarray(ev.meas_values[mid]),
        meas_cov=np.ascontiguousarray(ev.meas_cov[mid]),
        truth_start=np.ascontiguousarray(ev.truth_start[src]),
        truth_meas=np.ascontiguousarray(ev.truth_meas[mid]),
        meas_is_outlier=np.ascontiguousarray(ev.meas_is_outlier[mid]))
    return out, tuple(k for k, _ in cases)
