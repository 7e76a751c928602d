#!/usr/bin/env python3
"""Throughput of the batched Kalman track fit on one MI355X.

One step = one fit (forward filter + RTS smoother) of the whole synthetic
batch, inputs already resident in HBM.  Prints one JSON line.

Workload (default ``endcap``): 100,000 muons, 10 silicon barrel layers
(r = 32..1000 mm, half-length 3 m) plus 2x7 endcap discs (|z| = 600..2900 mm),
0.3 mm Si per surface, uniform Bz = 2 T, pT 1-10 GeV, |eta| < 2.5,
15/50 um measurements, plain FP64 fit.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = ("tracks/s fitted, mixed vs FP64 at stated pull accuracy; "
          "p50/p99 event latency")
FP64_VECTOR_PEAK_TFLOPS = 78.6  # MI355X published FP64 vector peak


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="endcap", choices=["endcap", "barrel"])
    ap.add_argument("--n-tracks", type=int, default=None)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--input-order", action="store_true",
                    help="fit tracks in input order instead of length-sorted")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run a plain device copy of this size (for the "
                         "HBM-counter calibration of rocprof runs)")
    return ap.parse_args()


def cpu_baseline(det, field, opts, events, threads):
    """Time the CPU oracle (restatement of the reference fitter) on the host."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle  # noqa: E402  (test infrastructure: the timed CPU baseline)
    oracle.load()
    n = events.n_tracks
    # warm-up on a small slice, then all tracks at `threads` threads
    oracle.fit(det, field, opts, events.subset(np.arange(min(n, 500))), n_threads=threads)
    _, secs = oracle.fit(det, field, opts, events, n_threads=threads)
    n1 = min(n, 5000)
    _, secs1 = oracle.fit(det, field, opts, events.subset(np.arange(n1)), n_threads=1)
    return {"value": n / secs, "unit": "tracks/s", "cores": threads, "kind": "port",
            "sample": f"all {n} tracks of the workload at {threads} threads "
                      f"(restated reference CPU fitter, g++ -O2)",
            "single_thread_value": n1 / secs1,
            "single_thread_sample": f"first {n1} tracks, 1 thread"}


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU (asked for --gpus {args.gpus}); "
              "tracks are independent, so N GPUs would each fit their own batch.",
              file=sys.stderr)
        return 2

    import torch
    import acts_amd as A
    from acts_amd.device import DeviceEvents, DeviceResult
    from acts_amd.flops import batch_flops

    det, field, events = A.workload(args.workload, n_tracks=args.n_tracks, seed=args.seed)
    opts = A.KalmanFitterOptions(length_sorted=not args.input_order)
    dev = torch.device("cuda:0")
    dev_in = DeviceEvents(events, dev)
    outputs = ("state_surface", "state_flags", "state_path", "state_chi2", "smoothed",
               "smoothed_cov")
    dev_out = DeviceResult(events.n_tracks, dev_in.n_slots, outputs, dev)
    fitter = A.KalmanFitter(det, field, opts)
    bin_, bout = dev_in.batch_in(), dev_out.batch_out()
    # a dedicated (default-priority) stream: its handle is non-null, so the
    # library launches on it and the torch events below time exactly the fits
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        fitter.fit_device(bin_, bout, stream.cuda_stream)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        fitter.fit_device(bin_, bout, stream.cuda_stream)
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = start.elapsed_time(end) / args.steps

    if args.copy_calibration_mib > 0:
        nbytes = args.copy_calibration_mib << 20
        src = torch.empty(nbytes // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()

    # per-track counters of the last fit -> algorithmic FLOP per launch
    class R:
        pass
    r = R()
    for name in ("status", "n_steps", "n_step_trials", "n_states", "n_measurements"):
        setattr(r, name, dev_out.host(name))
    flop = batch_flops(r)
    achieved = flop / (ms_per_step * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(REPO, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tpath) and args.n_tracks is None:
        with open(tpath) as f:
            tr = json.load(f)
        traffic = tr.get("bytes_per_launch")

    n = events.n_tracks
    line = {
        "metric": METRIC,
        "value": n / (ms_per_step * 1e-3),
        "unit": "tracks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (project generator, seed %d)" % args.seed,
        "config": {"workload": f"{args.workload}: {n} muons, "
                               f"{len(det.surfaces)} surfaces, Bz=2T, FP64 KF+RTS",
                   "n_tracks": n, "precision": "fp64",
                   "fit_failures": int(np.count_nonzero(r.status)),
                   "mean_states": float(np.mean(r.n_states)),
                   "mean_steps": float(np.mean(r.n_steps))},
        "roofline": {"bound": "valu", "achieved": achieved,
                     "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_VECTOR_PEAK_TFLOPS, "traffic": traffic,
                     "flop_per_launch": flop},
        "wall_s_timed": wall,
    }
    if not args.no_cpu_baseline:
        threads = max(1, min(16, os.cpu_count() or 1))
        line["cpu_baseline"] = cpu_baseline(det, field, opts, events, threads)
    fitter.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
