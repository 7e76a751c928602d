"""Event-serving loop: continuous batching of track fits on one GPU.

The reference processes events with the Examples' Sequencer, one
TrackFittingAlgorithm::execute per event, fitting the event's tracks one by
one (Examples/Algorithms/TrackFitting/src/TrackFittingAlgorithm.cpp:117-173).
Here events arrive as a Poisson stream; a dispatcher groups the events that
have arrived into batches, admitted by an HBM budget per batch, and hands
each batch to a free slot.  A slot is one fitter context (own stream,
staging buffers and scratch, geometry and field resident); its worker thread
copies the batch in, fits it, copies the results (surface, flags and
smoothed parameters of every state, track summaries) back, and stamps every event
of the batch with its completion time.  Latency of an event = completion -
arrival.
"""
from __future__ import annotations

import queue
import threading
import time
from dataclasses import dataclass

import numpy as np

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
    slots: int = 3                   # batches in flight
    hbm_budget_bytes: float = 4e9    # device memory one batch may use
    seed: int = 1


def event_stream(cfg: ServingConfig, pool_tracks: int):
    """(arrival times [s], multiplicities, offsets into the track pool)."""
    rng = np.random.default_rng(cfg.seed)
    n_max = int(cfg.rate_hz * cfg.duration_s * 2 + 16)
    gaps = rng.exponential(1.0 / cfg.rate_hz, size=n_max)
    t = np.cumsum(gaps)
    t = t[t < cfg.duration_s]
    m = np.exp(rng.normal(np.log(cfg.median_tracks), cfg.sigma, size=len(t)))
    m = np.clip(np.round(m), cfg.min_tracks, cfg.max_tracks).astype(np.int64)
    m = np.minimum(m, pool_tracks)
    off = rng.integers(0, pool_tracks - m + 1)
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
    """Device memory of one batch: inputs, scratch (744 B per state slot),
    serving outputs and per-track scratch."""
    per_track = 48 + 288 + 96 + 16 + 9 * 4 + 8
    per_slot = 744 + 4 + 4 + 48 + 4 + 52
    return n_tracks * (per_track + max_len * per_slot)


def serve(detector, field, pool: Events, cfg: ServingConfig,
          options: KalmanFitterOptions | None = None, device: int = 0) -> dict:
    """Run the serving loop; returns throughput and latency figures."""
    options = options or KalmanFitterOptions()
    arrivals, mult, offs = event_stream(cfg, pool.n_tracks)
    n_ev = len(arrivals)
    done_at = np.full(n_ev, np.nan)
    max_len = pool.max_sequence_length
    cap_tracks = max(cfg.max_tracks, int(cfg.hbm_budget_bytes // batch_bytes(1, max_len)))

    fitters = [KalmanFitter(detector, field, options, device=device) for _ in range(cfg.slots)]
    # warm each context up to the largest batch it may get (scratch sizes)
    warm = batch_of(pool, [(0, min(cap_tracks, pool.n_tracks))])
    for f in fitters:
        f.fit(warm, outputs=SERVING_OUTPUTS)

    free = queue.Queue()
    for i in range(cfg.slots):
        free.put(i)
    work = [queue.Queue() for _ in range(cfg.slots)]
    stats = {"batches": 0, "batch_tracks": []}
    t0 = [0.0]

    def worker(i):
        f = fitters[i]
        while True:
            item = work[i].get()
            if item is None:
                return
            ids = item
            ev = batch_of(pool, [(int(offs[e]), int(mult[e])) for e in ids])
            res = f.fit(ev, outputs=SERVING_OUTPUTS)
            now = time.perf_counter() - t0[0]
            done_at[ids] = now
            stats["batches"] += 1
            stats["batch_tracks"].append(ev.n_tracks)
            del res
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
            time.sleep(min(arrivals[nxt] - now, 0.002))
            continue
        slot = free.get()  # blocks while every slot is busy
        now = time.perf_counter() - t0[0]
        ids, tracks = [], 0
        while nxt < n_ev and arrivals[nxt] <= now and (
                not ids or tracks + mult[nxt] <= cap_tracks):
            ids.append(nxt)
            tracks += int(mult[nxt])
            nxt += 1
        work[slot].put(np.array(ids))
    for q in work:
        q.put(None)
    for th in threads:
        th.join()
    total = time.perf_counter() - t0[0]
    for f in fitters:
        f.close()
    lat = (done_at - arrivals) * 1e3
    return {"events": int(n_ev), "tracks": int(mult.sum()),
            "throughput_tracks_per_s": float(mult.sum() / total),
            "offered_tracks_per_s": float(mult.sum() / cfg.duration_s),
            "latency_ms_p50": float(np.percentile(lat, 50)),
            "latency_ms_p99": float(np.percentile(lat, 99)),
            "latency_ms_max": float(np.max(lat)),
            "batches": stats["batches"],
            "mean_batch_tracks": float(np.mean(stats["batch_tracks"])),
            "wall_s": float(total)}
