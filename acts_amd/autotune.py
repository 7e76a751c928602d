"""Autotuner for the scheduling knobs of the fit kernels.

The persistent forward kernel has five knobs that change only the speed of
a fit, never its results (acts_kf_options):
  sched_inner_steps  RK steps taken before the lanes waiting at a surface are
                     updated;
  sched_wait_64ths   the fraction of busy lanes that must wait at a surface
                     before the surface update runs early;
  length_sorted      the order in which tracks enter the queue: 1 longest
                     first, 0 input order, 2 by start direction, 3 by
                     length then momentum;
  sched_quarantine_steps  a wavefront holding a track older than this many
                     steps takes no new tracks until it ends;
  sched_transport_skip_steps  steps since the last surface after which a
                     track steps without the jacobian transport (helps
                     where loopers never reach another surface, as with
                     outliers that end at the loop-protection limit; hurts
                     where tracks do reach one, as in the field map).
`autotune` times candidate settings on a device-resident sample of the
workload by coordinate descent and keeps the fastest; results are cached in
a JSON file keyed by the workload's shape.

The kernels' shapes are not knobs: the forward kernel's register budget
(every lane holds its track's state, ~450 registers) allows one 64-lane
wavefront per SIMD, so a workgroup is one wavefront; forcing two per SIMD
spills, and running the forward pass as separate stepping and surface
kernels over the whole batch was measured slower (the lanes' state moves
through HBM every round).
"""
from __future__ import annotations

import dataclasses
import json
import os
import time
from typing import Callable

import numpy as np

SPACE = {
    "sched_inner_steps": [8, 16, 32, 64, 128],
    "sched_wait_64ths": [16, 32, 48, 64],
    "length_sorted": [1, 0, 2, 3],
    "sched_quarantine_steps": [0, 100, 200, 400],
    "sched_transport_skip_steps": [0, 64, 256],
}
DEFAULTS = {"sched_inner_steps": 64, "sched_wait_64ths": 64, "length_sorted": 1,
            "sched_quarantine_steps": 0, "sched_transport_skip_steps": 0}


def coordinate_descent(evaluate: Callable[[dict], float], space=SPACE, start=None,
                       rounds: int = 2) -> tuple[dict, dict]:
    """Minimise evaluate(config) one knob at a time; returns (best, {key: time})."""
    best = dict(start or DEFAULTS)
    seen: dict = {}

    def cost(cfg):
        key = json.dumps(cfg, sort_keys=True)
        if key not in seen:
            seen[key] = float(evaluate(dict(cfg)))
        return seen[key]

    best_t = cost(best)
    for _ in range(rounds):
        changed = False
        for knob, values in space.items():
            for v in values:
                cand = dict(best)
                cand[knob] = v
                t = cost(cand)
                if t < best_t:
                    best, best_t, changed = cand, t, True
        if not changed:
            break
    return best, seen


def signature(events, field, options, device_name: str = "") -> str:
    lens = np.diff(events.seq_offsets)
    n = events.n_tracks
    n_bucket = int(float(f"{n:.2g}"))
    return json.dumps({"n": n_bucket, "max_len": int(lens.max() if n else 0),
                       "mean_len": round(float(lens.mean()) if n else 0.0, 1),
                       "field": type(field).__name__, "precision": int(options.precision),
                       "device": device_name}, sort_keys=True)


def default_cache_path() -> str:
    base = os.environ.get("XDG_CACHE_HOME", os.path.join(os.path.expanduser("~"), ".cache"))
    return os.path.join(base, "acts_amd", "autotune.json")


def autotune(detector, field, events, options=None, steps: int = 3, warmup: int = 1,
             cache_path: str | None = None, device: int = 0,
             in_flight: int = 1) -> tuple[dict, dict]:
    """Tune the scheduling knobs on `events` (fitted on `device`); returns
    (best settings, measured milliseconds per setting).  Uses and updates the
    cache unless cache_path is the empty string.

    in_flight = 1 times one fit at a time; in_flight = k times the fits
    spread round-robin over k contexts, each on its own stream with its own
    outputs (the way an event loop keeps fits in flight), per fit.  The two
    can prefer different knobs: an order that starts the slowest tracks last
    leaves a long tail to one fit but fills it with the next fit's bulk."""
    import torch
    from .device import DeviceEvents, DeviceResult
    from .fitter import KalmanFitter, KalmanFitterOptions

    options = options or KalmanFitterOptions()
    dev = torch.device(f"cuda:{device}")
    sig = signature(events, field, options, torch.cuda.get_device_name(dev))
    if in_flight > 1:
        sig = json.dumps({"sig": sig, "in_flight": in_flight}, sort_keys=True)
    path = default_cache_path() if cache_path is None else cache_path
    cache = {}
    if path and os.path.exists(path):
        with open(path) as f:
            cache = json.load(f)
        if sig in cache:
            return cache[sig]["best"], cache[sig]["ms"]

    din = DeviceEvents(events, dev)
    douts = [DeviceResult(events.n_tracks, din.n_slots, ("smoothed",), dev)
             for _ in range(max(1, in_flight))]
    stream = torch.cuda.Stream(device=dev)

    def evaluate(cfg):
        opts = dataclasses.replace(options, **cfg)
        fitters = [KalmanFitter(detector, field, opts, device=device)
                   for _ in range(max(1, in_flight))]
        try:
            if in_flight <= 1:
                fitter, dout = fitters[0], douts[0]
                for _ in range(warmup):
                    fitter.fit_device(din.batch_in(), dout.batch_out(), stream.cuda_stream)
                torch.cuda.synchronize()
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                for _ in range(steps):
                    fitter.fit_device(din.batch_in(), dout.batch_out(), stream.cuda_stream)
                b.record(stream)
                torch.cuda.synchronize()
                return a.elapsed_time(b) / steps
            # the contexts' own streams (None), each fit with its own outputs
            for f, do in zip(fitters, douts):
                for _ in range(warmup):
                    f.fit_device(din.batch_in(), do.batch_out(), None)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for i in range(steps * in_flight):
                k = i % in_flight
                fitters[k].fit_device(din.batch_in(), douts[k].batch_out(), None)
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) * 1e3 / (steps * in_flight)
        finally:
            for f in fitters:
                f.close()

    best, seen = coordinate_descent(evaluate)
    ms = {k: v for k, v in seen.items()}
    if path:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        cache[sig] = {"best": best, "ms": ms}
        with open(path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return best, ms
