"""Edge-case batches for the fit tests, built from ordinary tracks: no
measurements, an empty surface sequence, a single measurement, a hole (a
measurement taken away, its surface kept), a measurement on a surface that
the track's sequence does not contain (never reached) together with the
same track without that measurement, measurements only on the last two
surfaces, and a second, displaced measurement on a surface that already
has one together with the same track without it, followed by ordinary
tracks.  `padded` moves a detector's surfaces to high indices."""
from __future__ import annotations

import dataclasses

import numpy as np

from acts_amd import capi
from acts_amd.detector import Detector, SurfaceDesc
from acts_amd.events import Events

# (kind, index of the source track)
CASES = (("no measurements", 0), ("empty sequence", 1), ("single measurement", 2),
         ("hole", 3), ("unreachable measurement", 4), ("unreachable removed", 4),
         ("last two measurements only", 5), ("duplicate measurement", 6),
         ("duplicate control", 6))
# loc0 displacement of the second measurement on a surface [mm]
DUPLICATE_SHIFT = 0.5


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
    if kind == "duplicate measurement":
        k = len(mi) // 2
        return seq, mi[:k + 1] + [mi[k]] + mi[k + 1:]
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
    for i, (kind, t) in enumerate(cases):
        if kind == "duplicate measurement":
            k = int(mo[t + 1] - mo[t]) // 2
            out.meas_values[new_mo[i] + k + 1, 0] += DUPLICATE_SHIFT
    return out, tuple(k for k, _ in cases)


def padded(det: Detector, ev: Events, n_pad: int) -> tuple[Detector, Events]:
    """The detector with `n_pad` unused cylinders (r = 5 m, no material)
    placed before its surfaces, and the events with their surface indices
    moved accordingly."""
    pad = [SurfaceDesc(capi.SURFACE_CYLINDER, np.eye(3), np.zeros(3), (5000.0, 3000.0), -1, False)
           for _ in range(n_pad)]
    d = Detector(surfaces=pad + list(det.surfaces), materials=list(det.materials),
                 name=det.name + "+pad")
    e = dataclasses.replace(ev, seq_surfaces=(ev.seq_surfaces + n_pad).astype(np.int32),
                            meas_surfaces=(ev.meas_surfaces + n_pad).astype(np.int32))
    return d, e


def shuffled(ev: Events, seed: int) -> Events:
    """The events with each track's measurements in a random order."""
    rng = np.random.default_rng(seed)
    mo = ev.meas_offsets
    perm = np.concatenate([mo[t] + rng.permutation(int(mo[t + 1] - mo[t]))
                           for t in range(ev.n_tracks)]).astype(np.int64)
    return dataclasses.replace(
        ev, meas_surfaces=np.ascontiguousarray(ev.meas_surfaces[perm]),
        meas_values=np.ascontiguousarray(ev.meas_values[perm]),
        meas_cov=np.ascontiguousarray(ev.meas_cov[perm]),
        truth_meas=np.ascontiguousarray(ev.truth_meas[perm]),
        meas_is_outlier=np.ascontiguousarray(ev.meas_is_outlier[perm]))
