"""Event-serving loop: continuous batching of track fits on one GPU.

The reference processes events with the Examples' Sequencer, one
TrackFittingAlgorithm::execute per event, fitting the event's tracks one by
one (Examples/Algorithms/TrackFitting/src/TrackFittingAlgorithm.cpp:117-173).
Here events arrive as a Poisson stream into a ring of pinned host memory
(the reader's buffers: each event's tracks follow the previous event's, and
an event that would run past the end of the ring starts again at its
beginning).  A dispatcher groups the events that have arrived into batches
of consecutive events, admitted by an HBM budget, and hands each batch to a
free slot.  A slot is one fitter context (own stream, staging buffers and
scratch, geometry and field resident) with pinned output buffers; its
worker thread fits the batch straight from the ring (only the two offset
arrays are rebased to the batch's first track, into pinned arrays of the
slot; the library's host-to-device copies read the ring by DMA), leaving
the results (surface, flags and smoothed parameters of every state, track
summaries) in pinned memory, and stamps every event of the batch with its
completion time.  Latency of an event = completion - arrival.

Admission: a batch is admitted when a slot is free and its device memory
(inputs, per-state scratch, outputs: `batch_bytes`) fits, together with the
batches already in flight, under the HBM budget; as batches retire their
bytes are returned and waiting events are admitted.  An event larger than
the whole budget runs alone.
"""
from __future__ import annotations

import ctypes as C
import queue
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import capi
from .events import Events
from .fitter import KalmanFitter, KalmanFitterOptions

# per-state results returned to the host for every event (the smoothed
# covariances stay on the device: they would triple the device-to-host
# traffic)
SERVING_OUTPUTS = ("state_surface", "state_flags", "smoothed")


@dataclass
class ServingConfig:
    rate_hz: float = 1000.0          # mean event arrival rate
    duration_s: float = 5.0          # length of the arrival window
    median_tracks: int = 5000        # lognormal multiplicity: median ...
    sigma: float = 0.7               # ... and width of log(multiplicity)
    min_tracks: int = 500
    max_tracks: int = 20000
    slots: int = 3                   # fitter contexts: most batches in flight
    hbm_budget_bytes: float = 32e9   # device memory of all batches in flight
    # tracks per batch: a batch's forward kernel lasts as long as its slowest
    # track (a looper that ends at max_steps: ~9 ms in the field map whether
    # the batch holds 48 000 tracks or 500 000), so large batches amortise
    # it; measured saturated capacity in the field map with 3 slots: 6.9 M
    # tracks/s at 60 000, 8.1 M at 150 000, 13.8 M at 300 000
    max_batch_tracks: int = 300000
    seed: int = 1
    keep_results: bool = False       # keep every event's results (tests)


def event_stream(cfg: ServingConfig, pool_tracks: int):
    """(arrival times [s], multiplicities, offsets into the track pool).

    The pool is the ring the events are read into: event e's tracks follow
    event e-1's, and an event that would run past the end of the pool
    starts at its beginning."""
    rng = np.random.default_rng(cfg.seed)
    n_max = int(cfg.rate_hz * cfg.duration_s * 2 + 16)
    gaps = rng.exponential(1.0 / cfg.rate_hz, size=n_max)
    t = np.cumsum(gaps)
    t = t[t < cfg.duration_s]
    m = np.exp(rng.normal(np.log(cfg.median_tracks), cfg.sigma, size=len(t)))
    m = np.clip(np.round(m), cfg.min_tracks, cfg.max_tracks).astype(np.int64)
    m = np.minimum(m, pool_tracks)
    off = np.zeros(len(t), np.int64)
    cursor = int(rng.integers(0, pool_tracks))
    for e in range(len(t)):
        if cursor + m[e] > pool_tracks:
            cursor = 0
        off[e] = cursor
        cursor += int(m[e])
    return t, m, off


def batch_of(pool: Events, ranges) -> Events:
    """The tracks pool[off:off+m] of every (off, m) in `ranges`, concatenated."""
    so, mo = pool.seq_offsets, pool.meas_offsets
    parts = {k: [] for k in ("start_params", "start_cov", "start_plane", "seq_surfaces",
                             "meas_surfaces", "meas_values", "meas_cov")}
    seq_len, meas_len = [], []
    for off, m in ranges:
        a, b = off, off + m
        parts["start_params"].append(pool.start_params[a:b])
        parts["start_cov"].append(pool.start_cov[a:b])
        parts["start_plane"].append(pool.start_plane[a:b])
        parts["seq_surfaces"].append(pool.seq_surfaces[so[a]:so[b]])
        parts["meas_surfaces"].append(pool.meas_surfaces[mo[a]:mo[b]])
        parts["meas_values"].append(pool.meas_values[mo[a]:mo[b]])
        parts["meas_cov"].append(pool.meas_cov[mo[a]:mo[b]])
        seq_len.append(np.diff(so[a:b + 1]))
        meas_len.append(np.diff(mo[a:b + 1]))
    n = int(sum(m for _, m in ranges))
    seq_off = np.zeros(n + 1, np.int64)
    seq_off[1:] = np.cumsum(np.concatenate(seq_len))
    meas_off = np.zeros(n + 1, np.int64)
    meas_off[1:] = np.cumsum(np.concatenate(meas_len))
    cat = {k: np.ascontiguousarray(np.concatenate(v)) for k, v in parts.items()}
    return Events(n_tracks=n, seq_offsets=seq_off, meas_offsets=meas_off,
                  truth_start=None, truth_meas=None, meas_is_outlier=None, **cat)


def batch_bytes(n_tracks: int, max_len: int) -> float:
    """Device memory of one batch: inputs, scratch (748 B per state slot),
    serving outputs and per-track scratch."""
    per_track = 48 + 288 + 96 + 16 + 9 * 4 + 8
    per_slot = 748 + 4 + 4 + 48 + 4 + 52
    return n_tracks * (per_track + max_len * per_slot)


def _pinned(n: int, w: int, dt) -> np.ndarray:
    """An uninitialised numpy array in pinned host memory."""
    import torch
    tdt = {np.dtype(np.float64): torch.float64, np.dtype(np.int32): torch.int32,
           np.dtype(np.int64): torch.int64, np.dtype(np.float32): torch.float32,
           np.dtype(np.uint8): torch.uint8}[np.dtype(dt)]
    shape = (n,) if w == 1 else (n, w)
    return torch.empty(shape, dtype=tdt, pin_memory=True).numpy()


RING_ARRAYS = ("start_params", "start_cov", "start_plane", "seq_offsets", "seq_surfaces",
               "meas_offsets", "meas_surfaces", "meas_values", "meas_cov", "meas_dims",
               "meas_indices")


def pinned_ring(pool: Events) -> Events:
    """The pool's fit inputs copied into pinned host memory: the ring the
    events are read into."""
    arrs = {}
    for name in RING_ARRAYS:
        a = getattr(pool, name)
        if a is None:
            arrs[name] = None
            continue
        a = np.ascontiguousarray(a)
        w = 1 if a.ndim == 1 else a.shape[1]
        arrs[name] = _pinned(a.shape[0], w, a.dtype)
        arrs[name][...] = a
    return Events(n_tracks=pool.n_tracks, truth_start=None, truth_meas=None,
                  meas_is_outlier=None, **arrs)


class _Slot:
    """One in-flight batch: a fitter context plus pinned host buffers for the
    batch's rebased offsets and its outputs, reused from batch to batch."""

    def __init__(self, fitter: KalmanFitter, cap_tracks: int, cap_seq: int):
        self.fitter = fitter
        self.seq_offsets = _pinned(cap_tracks + 1, 1, np.int64)
        self.meas_offsets = _pinned(cap_tracks + 1, 1, np.int64)
        self.out = {name: _pinned(cap_tracks, 1, np.int32)
                    for name in ("status", "n_states", "ndf", "n_measurements", "n_holes",
                                 "n_outliers", "n_steps", "n_step_trials")}
        self.out["chi2"] = _pinned(cap_tracks, 1, np.float32)
        self.out["state_surface"] = _pinned(cap_seq, 1, np.int32)
        self.out["state_flags"] = _pinned(cap_seq, 1, np.int32)
        self.out["smoothed"] = _pinned(cap_seq, 6, np.float64)

    def results(self, t0: int, cnt: int, s0: int, ns: int) -> dict:
        """Copies of the results of the batch's tracks t0..t0+cnt-1, whose
        state slots are s0..s0+ns-1."""
        out = {name: self.out[name][t0:t0 + cnt].copy()
               for name in ("status", "n_states", "chi2", "ndf", "n_measurements", "n_holes",
                            "n_outliers")}
        for name in SERVING_OUTPUTS:
            out[name] = self.out[name][s0:s0 + ns].copy()
        out["state_flags"] = out["state_flags"].view(np.uint32)
        return out

    def fit(self, ring: Events, lo: int, n: int) -> int:
        """Fit the tracks ring[lo:lo+n] in place (inputs read from the pinned
        ring, offsets rebased to the first track) and leave the results in
        the pinned outputs."""
        so, mo = ring.seq_offsets, ring.meas_offsets
        hi = lo + n
        s0, m0 = int(so[lo]), int(mo[lo])
        np.subtract(so[lo:hi + 1], s0, out=self.seq_offsets[:n + 1])
        np.subtract(mo[lo:hi + 1], m0, out=self.meas_offsets[:n + 1])
        bin_ = capi.BatchIn()
        bin_.n_tracks = n
        row = {"start_params": lo, "start_cov": lo, "start_plane": lo, "seq_surfaces": s0,
               "meas_surfaces": m0, "meas_values": m0, "meas_cov": m0, "meas_dims": m0,
               "meas_indices": m0}
        for name, r in row.items():
            a = getattr(ring, name)
            if a is not None:
                setattr(bin_, name, a.ctypes.data + r * a.strides[0])
        bin_.seq_offsets = self.seq_offsets.ctypes.data
        bin_.meas_offsets = self.meas_offsets.ctypes.data
        bin_.max_sequence_length = int(np.max(np.diff(self.seq_offsets[:n + 1]))) if n else 0
        bout = capi.BatchOut()
        for name in capi.BATCH_OUT_FIELDS:
            arr = self.out.get(name)
            setattr(bout, name, arr.ctypes.data if arr is not None else None)
        lib, ctx = self.fitter.lib, self.fitter.ctx
        t0 = time.perf_counter()
        capi.check(lib.acts_kf_fit_batch(ctx, C.byref(bin_), C.byref(bout), None),
                   "acts_kf_fit_batch")
        t1 = time.perf_counter()
        capi.check(lib.acts_kf_sync(ctx), "acts_kf_sync")
        self.last_ms = ((t1 - t0) * 1e3, (time.perf_counter() - t0) * 1e3)
        return n


def serve(detector, field, pool: Events, cfg: ServingConfig,
          options: KalmanFitterOptions | None = None, device: int = 0,
          ring: Events | None = None) -> dict:
    """Run the serving loop; returns throughput and latency figures (and,
    with cfg.keep_results, each event's results under "results" and its
    tracks as pool ranges under "event_ranges").  `ring`: the pool already
    in pinned memory (pinned_ring(pool)), to reuse it across runs."""
    options = options or KalmanFitterOptions()
    arrivals, mult, offs = event_stream(cfg, pool.n_tracks)
    n_ev = len(arrivals)
    done_at = np.full(n_ev, np.nan)
    max_len = max(pool.max_sequence_length, 1)
    cap_tracks = max(cfg.max_tracks,
                     min(cfg.max_batch_tracks, int(cfg.hbm_budget_bytes // batch_bytes(1, max_len))))

    if ring is None:
        ring = pinned_ring(pool)
    slots = [_Slot(KalmanFitter(detector, field, options, device=device), cap_tracks,
                   cap_tracks * max_len) for _ in range(cfg.slots)]
    # warm each context up to the largest batch it may get (scratch sizes)
    for sl in slots:
        sl.fit(ring, 0, min(cap_tracks, pool.n_tracks))

    free = queue.Queue()
    for i in range(cfg.slots):
        free.put(i)
    work = [queue.Queue() for _ in range(cfg.slots)]
    stats = {"batches": 0, "batch_tracks": [], "busy": 0, "max_busy": 0,
             "inflight": 0.0, "max_inflight": 0.0, "submit_ms": [], "fit_ms": []}
    results = {}
    cond = threading.Condition()
    t0 = [0.0]
    so, mo = pool.seq_offsets, pool.meas_offsets

    def worker(i):
        sl = slots[i]
        while True:
            item = work[i].get()
            if item is None:
                return
            ids, nbytes = item
            n = sl.fit(ring, int(offs[ids[0]]), int(mult[ids].sum()))
            done_at[ids] = time.perf_counter() - t0[0]
            if cfg.keep_results:
                t = s = 0
                for e in ids:
                    lo, cnt = int(offs[e]), int(mult[e])
                    ns = int(so[lo + cnt] - so[lo])
                    results[int(e)] = sl.results(t, cnt, s, ns)
                    t, s = t + cnt, s + ns
            with cond:
                stats["batches"] += 1
                stats["batch_tracks"].append(n)
                stats["submit_ms"].append(sl.last_ms[0])
                stats["fit_ms"].append(sl.last_ms[1])
                stats["busy"] -= 1
                stats["inflight"] -= nbytes
                cond.notify_all()
            free.put(i)

    threads = [threading.Thread(target=worker, args=(i,), daemon=True)
               for i in range(cfg.slots)]
    for th in threads:
        th.start()
    t0[0] = time.perf_counter()
    nxt = 0
    while nxt < n_ev:
        now = time.perf_counter() - t0[0]
        if arrivals[nxt] > now:
            time.sleep(min(arrivals[nxt] - now, 0.001))
            continue
        slot = free.get()  # blocks while every slot is busy
        with cond:
            while True:
                now = time.perf_counter() - t0[0]
                ids, tracks, nbytes = [], 0, 0.0
                while nxt + len(ids) < n_ev and arrivals[nxt + len(ids)] <= now:
                    e = nxt + len(ids)
                    b = batch_bytes(tracks + int(mult[e]), max_len)
                    if ids and tracks + mult[e] > cap_tracks:
                        break
                    # a batch is a contiguous range of the ring
                    if ids and offs[e] != offs[e - 1] + mult[e - 1]:
                        break
                    # admitted if it fits the budget left, or if nothing
                    # else is in flight (an event above the whole budget)
                    if stats["inflight"] + b > cfg.hbm_budget_bytes and (ids or stats["busy"]):
                        break
                    ids.append(e)
                    tracks += int(mult[e])
                    nbytes = b
                if ids:
                    break
                cond.wait()  # memory frees up as batches retire
            nxt += len(ids)
            stats["busy"] += 1
            stats["inflight"] += nbytes
            stats["max_busy"] = max(stats["max_busy"], stats["busy"])
            stats["max_inflight"] = max(stats["max_inflight"], stats["inflight"])
        work[slot].put((np.array(ids), nbytes))
    for q in work:
        q.put(None)
    for th in threads:
        th.join()
    total = time.perf_counter() - t0[0]
    for sl in slots:
        sl.fitter.close()
    lat = (done_at - arrivals) * 1e3
    rep = {"events": int(n_ev), "tracks": int(mult.sum()),
           "throughput_tracks_per_s": float(mult.sum() / total),
           "offered_tracks_per_s": float(mult.sum() / cfg.duration_s),
           "latency_ms_p50": float(np.percentile(lat, 50)),
           "latency_ms_p99": float(np.percentile(lat, 99)),
           "latency_ms_max": float(np.max(lat)),
           "batches": stats["batches"],
           "mean_batch_tracks": float(np.mean(stats["batch_tracks"])),
           # per batch: host time of the fit call (copies and launches
           # queued), and from the call to results in host memory
           "mean_batch_submit_ms": float(np.mean(stats["submit_ms"])),
           "mean_batch_fit_ms": float(np.mean(stats["fit_ms"])),
           "max_batches_in_flight": stats["max_busy"],
           "max_inflight_bytes": stats["max_inflight"],
           "hbm_budget_bytes": cfg.hbm_budget_bytes,
           "wall_s": float(total)}
    if cfg.keep_results:
        rep["results"] = results
        rep["event_ranges"] = [(int(o), int(m)) for o, m in zip(offs, mult)]
    return rep
