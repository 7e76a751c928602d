#!/usr/bin/env python3
"""Throughput of the batched Kalman track fit on one MI355X.

One step = one fit (forward filter + RTS smoother + track summary) of the
whole synthetic batch, inputs already resident in HBM.  The timed steps
alternate between --streams fitting contexts (default 3), each on its own
stream with its own outputs, so that several fits of the batch are in flight
and one's bulk overlaps the other's last long tracks, as in a production
loop; `single_stream` gives the one-fit-at-a-time figure and the kernel
timings are per launch on one context.  Prints one JSON line.

Workload (default `solenoid`): 500,000 muons, pT log-uniform 0.5-100 GeV,
|eta| < 2.5, through 10 silicon barrel cylinders (r = 32..1000 mm,
half-length 3 m) and 2x7 endcap discs (|z| = 600..2900 mm), 0.3 mm Si per
surface, in a 201x201x401 field map (x, y in +-2 m, z in +-4 m) sampled from
a 2 T solenoid of radius 1.2 m and length 6 m; 15/50 um measurements.
`value` is the plain FP64 fit (the drop-in); the mixed-precision fit (FP32
stepping and field map, FP64 Kalman algebra) of the same batch is timed
beside it with its error against the FP64 fit and its pulls.  Also in the
line: the fit with host<->device copies, the Examples' fit variants, the
track-state table, the 100,000-track endcap batch in a uniform field, the
2,000,000-track ragged batch under a hipGraph with autotuned scheduling,
the event-serving loop (capacity and p50/p99 latency, FP64 and mixed) and
the CPU baseline, whose results are compared with the GPU's track by track.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import platform
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = ("tracks/s fitted, mixed vs FP64 at stated pull accuracy; "
          "p50/p99 event latency")
FP64_VECTOR_PEAK_TFLOPS = 78.6  # MI355X published FP64 vector peak
HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
OUTPUTS = ("state_surface", "state_flags", "state_path", "state_chi2", "smoothed",
           "smoothed_cov")
# worker threads the box gives one GPU's process
CPU_THREADS_PER_GPU = 16


_T0 = time.perf_counter()


def log(msg: str) -> None:
    """Progress on stderr (the JSON line is the only stdout output)."""
    print(f"bench [{time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--streams", type=int, default=3,
                    help="fitting contexts, each on its own stream, that the timed "
                         "steps alternate between (1 = one fit at a time)")
    ap.add_argument("--workload", default="solenoid",
                    choices=["solenoid", "endcap", "barrel", "ragged"])
    ap.add_argument("--n-tracks", type=int, default=None)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the perigee / MBF / Sympy / Examples' variants and the "
                         "track-state table")
    ap.add_argument("--no-endcap", action="store_true",
                    help="skip the 100,000-track uniform-field endcap batch")
    ap.add_argument("--ragged-tracks", type=int, default=2_000_000,
                    help="tracks of the ragged batch under a hipGraph (0 = skip)")
    ap.add_argument("--input-order", action="store_true",
                    help="queue tracks in input order instead of longest first")
    ap.add_argument("--serving-seconds", type=float, default=2.0,
                    help="length of each loaded event-serving run (0 = skip serving)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="time the library's default scheduling knobs instead of tuning "
                         "them (per precision, on the batch itself)")
    ap.add_argument("--scheduling", default="",
                    help="scheduling knobs as JSON instead of tuning them: one set for both "
                         "precisions, or {\"fp64\": {...}, \"mixed\": {...}} as the "
                         "output's config.scheduling prints them")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run a plain device copy of this size (for the "
                         "HBM-counter calibration of rocprof runs)")
    return ap.parse_args()


def host_cpu() -> dict:
    """The host's CPU model and core counts."""
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return {"model": model, "os_cpu_count": os.cpu_count(), "affinity_cpus": affinity}


def cpu_baseline(det, field, opts, events, threads):
    """Time the CPU oracle (restatement of the reference fitter) on the host
    cores: the whole batch at `threads` threads (its results are returned
    for the parity check) and a 5,000-track sample on one thread."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle  # noqa: E402  (test infrastructure: the timed CPU baseline)
    oracle.load()
    n = events.n_tracks
    oracle.fit(det, field, opts, events.subset(np.arange(min(n, 500))), n_threads=threads)
    ref, secs = oracle.fit(det, field, opts, events, n_threads=threads,
                           outputs=("state_surface", "state_flags", "smoothed", "smoothed_cov"))
    n1 = min(n, 5000)
    _, secs1 = oracle.fit(det, field, opts, events.subset(np.arange(n1)), n_threads=1)
    cpu = host_cpu()
    line = {"value": n / secs, "unit": "tracks/s", "cores": threads, "kind": "port",
            "sample": f"the whole batch ({n} tracks) at {threads} threads "
                      f"(restated reference CPU fitter, g++ -O2, FP64), {secs:.1f} s",
            "single_thread_value": n1 / secs1,
            "single_thread_sample": f"first {n1} tracks, 1 thread, {secs1:.1f} s",
            "cpu_model": cpu["model"], "os_cpu_count": cpu["os_cpu_count"],
            "affinity_cpus": cpu["affinity_cpus"],
            "threads_note": f"{CPU_THREADS_PER_GPU} worker threads: the CPU share the GPU "
                            "host gives one GPU's process"}
    return line, ref


def serving(args, det, field, pool):
    """Event serving (tracks as in the field-map workload): Poisson arrivals,
    lognormal multiplicity (median 5000, clipped 500-20000).  A saturated
    run (every event arrived at once, >= 1 s of fitting) measures the
    loop's capacity for each precision; events are then offered at 0.5, 0.8
    and 0.95 of the FP64 capacity, to both precisions."""
    import acts_amd as A
    from acts_amd import capi
    from acts_amd.serving import ServingConfig, pinned_ring, serve
    ring = pinned_ring(pool)  # the reader's pinned buffers, shared by the runs
    mean_tracks = 5000 * np.exp(0.7 ** 2 / 2)
    precs = (("fp64", capi.PRECISION_FP64), ("mixed", capi.PRECISION_MIXED))
    base = ServingConfig(seed=args.seed)
    # the field map's scheduling knobs for batches of events: longest
    # sequence first (queue orders that start slow tracks last lengthen each
    # batch), quarantine and inner steps as tuned on the field-map batch
    # (measured capacity 13.7 M tracks/s FP64 against 12.8 M with the
    # library's defaults and 13.1 M with queue order 3)
    knobs = dict(length_sorted=1, sched_quarantine_steps=100, sched_inner_steps=16)
    out = {"config": {"multiplicity": "lognormal median 5000 sigma 0.7 clipped 500-20000",
                      "slots": base.slots, "hbm_budget_bytes": base.hbm_budget_bytes,
                      "max_batch_tracks": base.max_batch_tracks,
                      "pool_tracks": pool.n_tracks,
                      "tracks": "pT 0.5-100 GeV log-uniform, |eta|<2.5, 201x201x401 "
                                "solenoid map", "latency": "arrival -> results on host",
                      "duration_s": args.serving_seconds, "scheduling": knobs}}
    cap = {}
    for name, prec in precs:
        # enough events for >= 1 s at ~15 M tracks/s, all arrived at once
        sat_cfg = dataclasses.replace(base, rate_hz=4e4, duration_s=0.1)
        sat = serve(det, field, pool, sat_cfg, A.KalmanFitterOptions(precision=prec, **knobs),
                    ring=ring)
        cap[name] = sat["throughput_tracks_per_s"]
        log(f"serving {name}: capacity {cap[name]:.3g} tracks/s")
        out[f"saturated_{name}"] = sat
    out["capacity_tracks_per_s"] = cap
    out["mixed_over_fp64_capacity"] = cap["mixed"] / cap["fp64"]
    for load in (0.5, 0.8, 0.95):
        rate = load * cap["fp64"] / mean_tracks
        cfg = dataclasses.replace(base, rate_hz=rate, duration_s=args.serving_seconds)
        entry = {"rate_hz": rate}
        for name, prec in precs:
            r = serve(det, field, pool, cfg, A.KalmanFitterOptions(precision=prec, **knobs),
                      ring=ring)
            entry[name] = {k: r[k] for k in ("throughput_tracks_per_s", "latency_ms_p50",
                                             "latency_ms_p99", "latency_ms_max", "events",
                                             "batches", "mean_batch_tracks",
                                             "max_batches_in_flight")}
        out[f"load_{load}"] = entry
    return out


def timed_fits(fitter, bin_, bout, stream, steps, warmup):
    import torch
    for _ in range(warmup):
        fitter.fit_device(bin_, bout, stream.cuda_stream)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    fwd = []
    for _ in range(steps):
        fitter.fit_device(bin_, bout, stream.cuda_stream)
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # kernel durations of the forward filter and the smoother, measured with
    # HIP events on the stream the kernels run on (one extra fit per sample)
    if not fitter.options.use_graph:  # a replayed graph records no events
        for _ in range(min(steps, 5)):
            fitter.fit_device(bin_, bout, stream.cuda_stream)
            fwd.append(fitter.kernel_times())
    return start.elapsed_time(end) / steps, wall, fwd


def timed_overlapped(fitters, bin_, bouts, steps, warmup):
    """Time `steps` fits of the batch spread round-robin over several
    contexts, each on the context's own stream and with its own outputs, so
    that one fit's bulk runs while another's last long tracks finish.
    Returns ms per fit over the timed region, bracketed by device
    synchronisations and timed by the host clock.  (The contexts' own
    streams run concurrently; two streams of torch's pool were measured to
    serialise the fits.)"""
    import torch
    for f, bo in zip(fitters, bouts):
        for _ in range(warmup):
            f.fit_device(bin_, bo, None)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        k = i % len(fitters)
        fitters[k].fit_device(bin_, bouts[k], None)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def host_copies(det, field, opts, events, steps):
    """The fit through the host-buffer entry point (acts_kf_fit_batch):
    inputs copied host->device, outputs (surface, flags, path, chi2,
    smoothed parameters and covariance of every state, the summaries)
    device->host, from and to pinned host memory; one fit at a time, wall
    clock from the call to acts_kf_sync."""
    import ctypes as C
    import torch
    import acts_amd as A
    from acts_amd import capi
    from acts_amd.fitter import STATE_ARRAYS, TRACK_ARRAYS

    def pinned_like(a):
        t = torch.empty(a.shape, dtype={np.float64: torch.float64, np.int64: torch.int64,
                                         np.int32: torch.int32}[a.dtype.type], pin_memory=True)
        out = t.numpy()
        out[...] = a
        return t, out

    keep = []
    bin_ = capi.BatchIn()
    bin_.n_tracks = events.n_tracks
    in_bytes = 0
    for name in ("start_params", "start_cov", "start_plane", "seq_offsets", "seq_surfaces",
                 "meas_offsets", "meas_surfaces", "meas_values", "meas_cov"):
        t, arr = pinned_like(getattr(events, name))
        keep.append(t)
        in_bytes += arr.nbytes
        setattr(bin_, name, arr.ctypes.data)
    bin_.max_sequence_length = events.max_sequence_length
    n, ns = events.n_tracks, int(events.seq_offsets[-1])
    bout = capi.BatchOut()
    out_bytes = 0
    for name, dt in TRACK_ARRAYS.items():
        t, arr = pinned_like(np.zeros(n, dt if dt != np.float32 else np.int32))
        keep.append(t)
        out_bytes += arr.nbytes
        setattr(bout, name, arr.ctypes.data)
    for name in OUTPUTS:
        dt, w = STATE_ARRAYS[name]
        dt = {np.float32: np.int32, np.uint32: np.int32}.get(dt, dt)
        t, arr = pinned_like(np.zeros((ns, w) if w > 1 else ns, dt))
        keep.append(t)
        out_bytes += arr.nbytes
        setattr(bout, name, arr.ctypes.data)
    fitter = A.KalmanFitter(det, field, opts)
    lib, ctx = fitter.lib, fitter.ctx
    times = []
    try:
        for i in range(steps + 2):
            t0 = time.perf_counter()
            capi.check(lib.acts_kf_fit_batch(ctx, C.byref(bin_), C.byref(bout), None),
                       "acts_kf_fit_batch")
            capi.check(lib.acts_kf_sync(ctx), "acts_kf_sync")
            if i >= 2:
                times.append(time.perf_counter() - t0)
    finally:
        fitter.close()
    ms = float(np.median(times)) * 1e3
    return {"value": n / (ms * 1e-3), "ms_per_fit": ms, "bytes_to_device": in_bytes,
            "bytes_from_device": out_bytes, "host_memory": "pinned",
            "note": "acts_kf_fit_batch + acts_kf_sync, one fit at a time, median of "
                    f"{steps}; includes the PCIe copies of inputs and outputs"}


# algorithmic bytes of the track-state table kernel: every slot writes its
# row (154 float + 4 int columns) and reads its track lookup; a state reads
# three parameter vectors and covariances (1008 B) and its surface, flags,
# path, chi2 and measurement index (24 B); a state with a measurement also
# reads the measurement (48 B) and the truth (48 B)
TS_SLOT_BYTES = 154 * 4 + 4 * 4 + 16
TS_STATE_BYTES = 3 * (6 + 36) * 8 + 24
TS_LINK_BYTES = 48 + 48


def track_states_bench(det, field, opts, events, dev_in, stream, steps):
    """Time the track-state table kernel (RootTrackStatesWriter's columns)
    on one fitted batch: HBM-bound, priced against the HBM peak."""
    import torch
    import acts_amd as A
    from acts_amd import capi
    from acts_amd.device import DeviceResult, to_device
    outs = ("state_surface", "state_flags", "state_path", "state_chi2", "state_measurement",
            "predicted", "predicted_cov", "filtered", "filtered_cov", "smoothed",
            "smoothed_cov")
    dev = stream.device
    res = DeviceResult(events.n_tracks, dev_in.n_slots, outs, dev)
    fitter = A.KalmanFitter(det, field, opts)
    fitter.fit_device(dev_in.batch_in(), res.batch_out(), stream.cuda_stream)
    truth = to_device(events.truth_meas, dev)
    floats = torch.empty((capi.TS_NFLOAT, dev_in.n_slots), dtype=torch.float32, device=dev)
    ints = torch.empty((capi.TS_NINT, dev_in.n_slots), dtype=torch.int32, device=dev)
    o = capi.TrackStatesOut()
    o.n_slots, o.columns, o.ints = dev_in.n_slots, floats.data_ptr(), ints.data_ptr()
    bin_, bout = dev_in.batch_in(), res.batch_out()

    def launch():
        fitter.track_states_device(bin_, bout, truth.data_ptr(), o, stream.cuda_stream)
    for _ in range(3):
        launch()
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        launch()
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / steps
    fitter.close()
    st = ints[capi.TS_STATE_TYPE].cpu().numpy()
    n_states = int(np.count_nonzero(st >= 0))
    n_link = int(np.count_nonzero(ints[capi.TS_DIM_HIT].cpu().numpy() > 0))
    nbytes = (dev_in.n_slots * TS_SLOT_BYTES + n_states * TS_STATE_BYTES
              + n_link * TS_LINK_BYTES)
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"kernel": "trackStatesKernel", "ms": ms, "states": n_states,
            "states_per_s": n_states / (ms * 1e-3), "bytes_per_launch": nbytes,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS}}


class Bench:
    """Timing of one workload's fits on the device-resident batch."""

    def __init__(self, args, det, field, events, wl_opts, tuned):
        """`tuned`: scheduling knobs, one dict for every precision or a dict
        of them keyed by precision."""
        import torch
        from acts_amd.device import DeviceEvents
        self.args = args
        self.det, self.field, self.events = det, field, events
        self.dev = torch.device("cuda:0")
        self.dev_in = DeviceEvents(events, self.dev)
        self.bin = self.dev_in.batch_in()
        self.stream = torch.cuda.Stream(device=self.dev)  # default priority
        self.wl_opts, self.tuned = wl_opts, tuned
        self.single = {}

    def run(self, precision, extra=None, outputs=OUTPUTS, streams=None, steps=None):
        """ms per fit with `streams` fits in flight, the wall time, per-launch
        kernel times of context 0 and context 0's outputs; the one-at-a-time
        figure is kept in self.single."""
        import acts_amd as A
        from acts_amd.device import DeviceResult
        streams = streams or self.args.streams
        steps = steps or self.args.steps
        n = self.events.n_tracks
        tuned = self.tuned.get(precision, {}) if self.tuned and all(
            isinstance(v, dict) for v in self.tuned.values()) else self.tuned
        opts = dataclasses.replace(self.wl_opts, precision=precision, kernel_timing=True,
                                   **tuned, **(extra or {}))
        outs = [DeviceResult(n, self.dev_in.n_slots, outputs, self.dev) for _ in range(streams)]
        fitters = [A.KalmanFitter(self.det, self.field, opts) for _ in range(streams)]
        ms1, wall, kt = timed_fits(fitters[0], self.bin, outs[0].batch_out(), self.stream,
                                   steps, self.args.warmup)
        ms = ms1
        if streams > 1:
            ms = timed_overlapped(fitters, self.bin, [o.batch_out() for o in outs], steps,
                                  self.args.warmup)
        self.single[(precision, tuple(sorted((extra or {}).items())))] = ms1
        for f in fitters:
            f.close()
        return ms, wall, kt, outs[0]


def same_surface_tracks(sa, oa, na, sb, ob, nb) -> np.ndarray:
    """Per track, whether two fits' state-surface sequences are equal
    (surfaces sa/sb in slot layouts with offsets oa/ob, state counts na/nb)."""
    eq = na == nb
    n = np.where(eq, na, 0).astype(np.int64)
    t = np.repeat(np.arange(len(n)), n)
    k = np.arange(len(t)) - np.repeat(np.cumsum(n) - n, n)
    bad = np.bincount(t, weights=(sa[oa[t] + k] != sb[ob[t] + k]), minlength=len(n))
    return eq & (bad == 0)


def summaries(res):
    class R:
        pass
    r = R()
    for name in ("status", "n_steps", "n_step_trials", "n_states", "n_measurements"):
        setattr(r, name, res.host(name))
    return r


def endcap_section(args):
    """The 100,000-track endcap batch in the uniform 2 T field: FP64 and
    mixed, fits in flight and one at a time."""
    import acts_amd as A
    from acts_amd import capi
    from acts_amd import quality
    from acts_amd.flops import forward_flops
    det, field, events = A.workload("endcap", seed=args.seed)
    b = Bench(args, det, field, events, A.workload_options("endcap"), {})
    n = events.n_tracks
    ms, _, kt, out = b.run(capi.PRECISION_FP64)
    fwd_ms = float(np.mean([k[0] for k in kt]))
    flop = forward_flops(summaries(out))
    first64 = quality.first_states(out, events)
    line = {"workload": f"endcap: {n} muons, {len(det.surfaces)} surfaces, uniform Bz = 2 T, "
                        "pT 1-10 GeV, |eta| < 2.5",
            "fp64": {"value": n / (ms * 1e-3), "ms_per_step": ms,
                     "single_stream_value": n / (b.single[(capi.PRECISION_FP64, ())] * 1e-3),
                     "forward_kernel_ms": fwd_ms,
                     "smoother_ms": float(np.mean([k[1] for k in kt])),
                     "forward_frac_of_fp64_peak": flop / (fwd_ms * 1e-3) / 1e12
                     / FP64_VECTOR_PEAK_TFLOPS},
            "pulls": quality.pulls(first64, events)}
    del out
    if not args.no_mixed:
        msm, _, ktm, outm = b.run(capi.PRECISION_MIXED)
        firstm = quality.first_states(outm, events)
        line["mixed"] = {"value": n / (msm * 1e-3), "ms_per_step": msm,
                         "single_stream_value":
                             n / (b.single[(capi.PRECISION_MIXED, ())] * 1e-3),
                         "forward_kernel_ms": float(np.mean([k[0] for k in ktm])),
                         "first_state_smoothed_rel_err_vs_fp64":
                             quality.relative_error(firstm, first64)}
    return line


def ragged_section(args):
    """The ragged batch (6-30 measurements, holes, outliers) replayed from a
    captured hipGraph with scheduling knobs chosen by the autotuner on the
    batch itself with two fits in flight (FP64; the mixed fit uses the same
    knobs: tuned separately they came out nearly the same, at twice the
    tuning time), FP64 and mixed."""
    import acts_amd as A
    from acts_amd import capi
    from acts_amd.autotune import autotune
    log(f"ragged: generating {args.ragged_tracks} tracks")
    det, field, events = A.workload("ragged", n_tracks=args.ragged_tracks, seed=args.seed)
    opts = A.workload_options("ragged")
    lens = np.diff(events.meas_offsets)
    line = {"workload": f"ragged: {events.n_tracks} muons, {len(det.surfaces)} surfaces, "
                        f"{int(lens.min())}-{int(lens.max())} measurements "
                        f"(mean {lens.mean():.1f}), 5 % holes, 10 % of tracks with 1-3 "
                        "outliers at 20 sigma (chi2 cut 25), uniform Bz = 2 T"}
    t0 = time.perf_counter()
    tuned = None
    for name, prec in (("fp64", capi.PRECISION_FP64), ("mixed", capi.PRECISION_MIXED)):
        if name == "mixed" and args.no_mixed:
            continue
        o = dataclasses.replace(opts, precision=prec)
        if tuned is None:
            tuned, _ = autotune(det, field, events, o, cache_path="", in_flight=2, steps=2)
            log(f"ragged {name}: tuned {tuned}")
        b = Bench(args, det, field, events, o, dict(tuned, use_graph=True))
        ms, _, _, _ = b.run(prec, streams=2, steps=max(3, args.steps // 4))
        ms1 = b.single[(prec, ())]
        line[name] = {"value": events.n_tracks / (ms * 1e-3), "ms_per_step": ms,
                      "single_stream_value": events.n_tracks / (ms1 * 1e-3),
                      "streams": 2, "scheduling": tuned, "use_graph": True}
    line["autotune_and_timing_s"] = time.perf_counter() - t0
    return line


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU (asked for --gpus {args.gpus}); "
              "tracks are independent, so N GPUs would each fit their own batch.",
              file=sys.stderr)
        return 2

    import torch
    import acts_amd as A
    from acts_amd import capi
    from acts_amd import quality
    from acts_amd.flops import batch_flops, forward_flops

    t_start = time.perf_counter()
    log(f"generating the {args.workload} workload")
    det, field, events = A.workload(args.workload, n_tracks=args.n_tracks, seed=args.seed)
    n = events.n_tracks
    log(f"{n} tracks generated")
    wl_opts = A.workload_options(args.workload)
    tuned = {}
    autotuned = False
    if args.scheduling:
        tuned = json.loads(args.scheduling)
        if "fp64" in tuned:
            tuned = {capi.PRECISION_FP64: tuned["fp64"],
                     capi.PRECISION_MIXED: tuned.get("mixed", tuned["fp64"])}
    elif args.input_order:
        tuned = {"length_sorted": 0}
    elif args.autotune:
        # the autotuner's coordinate descent over the scheduling knobs, per
        # precision, on the batch itself (a 100,000-track sample of the
        # field-map batch once chose knobs 30 % slower on the whole batch
        # than the defaults: the tail of a small batch behaves differently),
        # timed with as many fits in flight as the timed steps keep; the
        # timed fits use them
        from acts_amd.autotune import autotune
        t0 = time.perf_counter()
        tuned = {}
        for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
            tuned[prec], _ = autotune(det, field, events,
                                      dataclasses.replace(wl_opts, precision=prec),
                                      cache_path="", in_flight=args.streams)
        autotuned = True
        log(f"autotuned in {time.perf_counter() - t0:.1f} s: {tuned}")
    b = Bench(args, det, field, events, wl_opts, tuned)

    ms64, wall64, kt64, out64 = b.run(capi.PRECISION_FP64)
    torch.cuda.synchronize()
    log(f"FP64: {ms64:.3f} ms per fit in flight")

    if args.copy_calibration_mib > 0:
        nbytes = args.copy_calibration_mib << 20
        src = torch.empty(nbytes // 8, dtype=torch.float64, device=b.dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()

    r = summaries(out64)
    flop_fwd = forward_flops(r)
    flop_all = batch_flops(r)
    fwd_ms = float(np.mean([k[0] for k in kt64]))
    smooth_ms = float(np.mean([k[1] for k in kt64]))
    achieved = flop_fwd / (fwd_ms * 1e-3) / 1e12
    in_flight = flop_fwd / (ms64 * 1e-3) / 1e12
    traffic, traffic_src = None, None
    tpath = os.path.join(REPO, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tpath) and args.n_tracks is None:
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
        traffic_src = (f"profiles/traffic_{args.workload}.json: rocprofv3 --pmc FETCH_SIZE "
                       "and WRITE_SIZE of bench.py on this workload (commands there), scaled "
                       "by a plain device copy; not measured in this run")

    field_desc = ("201x201x401 solenoid field map" if args.workload == "solenoid"
                  else "uniform Bz = 2 T")
    line = {
        "metric": METRIC,
        "value": n / (ms64 * 1e-3),
        "unit": "tracks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (project generator, seed %d)" % args.seed,
        "config": {"workload": f"{args.workload}: {n} muons, {len(det.surfaces)} surfaces, "
                               f"{field_desc}, FP64 KF + RTS smoother",
                   "n_tracks": n,
                   "fit_failures": int(np.count_nonzero(r.status)),
                   "mean_states": float(np.mean(r.n_states)),
                   "mean_steps": float(np.mean(r.n_steps)),
                   "streams": args.streams},
        "roofline": {"bound": "valu", "kernel": "fitForwardKernel (FP64)",
                     "achieved": achieved, "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_VECTOR_PEAK_TFLOPS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "flop_per_launch": flop_fwd, "kernel_ms": fwd_ms,
                     "frac_note": "frac = algorithmic FLOP of one forward launch / its "
                                  "duration (one fit alone); frac_in_flight = the same FLOP "
                                  "/ ms_per_step with the fits in flight",
                     "frac_in_flight": in_flight / FP64_VECTOR_PEAK_TFLOPS,
                     "smoother_ms": smooth_ms, "fit_flop_per_step": flop_all},
    }
    ms1 = b.single[(capi.PRECISION_FP64, ())]
    line["single_stream"] = {"value": n / (ms1 * 1e-3), "ms_per_step": ms1,
                             "note": "one fit at a time on one stream"}
    first64 = quality.first_states(out64, events)
    line["pulls"] = quality.pulls(first64, events)

    if not args.no_mixed:
        msm, _, ktm, outm = b.run(capi.PRECISION_MIXED)
        firstm = quality.first_states(outm, events)
        fwdm = float(np.mean([k[0] for k in ktm]))
        line["mixed"] = {
            "value": n / (msm * 1e-3), "ms_per_step": msm,
            "dtype": "f32 stepping and field map, f64 Kalman algebra",
            "over_fp64": msm and ms64 / msm,
            "single_stream_value": n / (b.single[(capi.PRECISION_MIXED, ())] * 1e-3),
            "forward_kernel_ms": fwdm,
            "smoother_ms": float(np.mean([k[1] for k in ktm])),
            "pulls": quality.pulls(firstm, events),
            "first_state_smoothed_rel_err_vs_fp64": quality.relative_error(firstm, first64)}
        del outm
        log(f"mixed: {msm:.3f} ms per fit in flight")
    knobs64 = tuned.get(capi.PRECISION_FP64, {}) if capi.PRECISION_FP64 in tuned else tuned
    line["with_host_copies"] = host_copies(det, field, dataclasses.replace(wl_opts, **knobs64),
                                           events, 5)
    log("host copies done")

    if not args.no_variants:
        msp, _, _, outp = b.run(capi.PRECISION_FP64,
                                {"reference_surface": capi.REFERENCE_PERIGEE,
                                 "reference_strategy": capi.EXTRAPOLATE_FIRST},
                                OUTPUTS + ("reference_params", "reference_cov"))
        line["with_perigee"] = {
            "value": n / (msp * 1e-3), "ms_per_step": msp,
            "extrapolation_failures": int(np.count_nonzero(outp.host("status")) -
                                          np.count_nonzero(r.status))}
        del outp
        line["track_states"] = track_states_bench(det, field, wl_opts, events, b.dev_in,
                                                  b.stream, args.steps)
        msb, _, ktb, outb = b.run(capi.PRECISION_FP64, {"smoother": capi.SMOOTHER_MBF})
        line["with_mbf"] = {
            "value": n / (msb * 1e-3), "ms_per_step": msb,
            "smoother_ms": float(np.mean([k[1] for k in ktb])),
            "first_state_rel_err_vs_rts": quality.relative_error(
                quality.first_states(outb, events), first64)}
        del outb
        msr, _, ktr, outr = b.run(capi.PRECISION_FP64, {"reverse_filtering": True})
        line["with_reverse_filtering"] = {
            "value": n / (msr * 1e-3), "ms_per_step": msr,
            "note": "every track filtered in reverse instead of smoothed "
                    "(KalmanFitterOptions::reverseFiltering)",
            "reversed_tracks": int(np.count_nonzero(outr.host("reversed"))),
            "fit_failures": int(np.count_nonzero(outr.host("status")))}
        del outr
        mss, _, kts, outs_ = b.run(capi.PRECISION_FP64, {"stepper": capi.STEPPER_SYMPY})
        line["with_sympy_stepper"] = {
            "value": n / (mss * 1e-3), "ms_per_step": mss,
            "forward_kernel_ms": float(np.mean([k[0] for k in kts])),
            "first_state_rel_err_vs_eigen": quality.relative_error(
                quality.first_states(outs_, events), first64)}
        del outs_
        # the layer navigator: surfaces found from the geometry instead of
        # the given sequences, each track with 6 spare state slots
        ev_l = A.with_state_slots(events, np.diff(events.seq_offsets) + 6)
        bl = Bench(args, det, field, ev_l, wl_opts, tuned)
        msl, _, ktl, outl = bl.run(capi.PRECISION_FP64, {"navigator": capi.NAVIGATOR_LAYERS},
                                   ("state_surface",))
        st_l = outl.host("status")
        same = same_surface_tracks(out64.host("state_surface"), events.seq_offsets, r.n_states,
                                   outl.host("state_surface"), ev_l.seq_offsets,
                                   outl.host("n_states"))
        both = (r.status == 0) & (st_l == 0)
        line["with_layer_navigator"] = {
            "value": n / (msl * 1e-3), "ms_per_step": msl,
            "forward_kernel_ms": float(np.mean([k[0] for k in ktl])),
            "note": "surfaces found from the geometry (acts_kf_options.navigator = LAYERS) "
                    "instead of the given sequences",
            "fit_failures": int(np.count_nonzero(st_l)),
            "same_surfaces_as_direct": float(np.count_nonzero(same & both) /
                                             max(np.count_nonzero(both), 1))}
        del outl, bl
        if not args.no_mixed:
            msx, _, ktx, outx = b.run(capi.PRECISION_MIXED, {"stepper": capi.STEPPER_SYMPY})
            line["with_sympy_stepper"]["mixed"] = {
                "value": n / (msx * 1e-3), "ms_per_step": msx,
                "forward_kernel_ms": float(np.mean([k[0] for k in ktx])),
                "first_state_rel_err_vs_fp64_eigen": quality.relative_error(
                    quality.first_states(outx, events), first64)}
            del outx
        mse, _, kte, oute = b.run(capi.PRECISION_FP64,
                                  {"stepper": capi.STEPPER_SYMPY,
                                   "smoother": capi.SMOOTHER_MBF,
                                   "reference_surface": capi.REFERENCE_PERIGEE,
                                   "reference_strategy": capi.EXTRAPOLATE_FIRST},
                                  OUTPUTS + ("reference_params", "reference_cov"))
        line["examples_configuration"] = {
            "value": n / (mse * 1e-3), "ms_per_step": mse,
            "forward_kernel_ms": float(np.mean([k[0] for k in kte])),
            "config": "SympyStepper + MbfSmoother + perigee (first measurement)",
            "fit_failures": int(np.count_nonzero(oute.host("status")))}
        del oute
        log("variants done")

    if tuned and capi.PRECISION_FP64 not in tuned:
        line["config"]["scheduling"] = tuned
    elif tuned:
        line["config"]["scheduling"] = {
            "fp64": tuned[capi.PRECISION_FP64], "mixed": tuned[capi.PRECISION_MIXED],
            "tuned_on": (f"the batch ({n} tracks, acts_amd.autotune, {args.streams} fits "
                         "in flight as in the timed steps)"
                         if autotuned else "given (--scheduling)")}
    if autotuned:
        b.tuned = {}
        msd, _, _, _ = b.run(capi.PRECISION_FP64)
        b.tuned = tuned
        line["default_scheduling"] = {
            "value": n / (msd * 1e-3), "ms_per_step": msd,
            "note": "FP64 with the library's default scheduling knobs"}

    if not args.no_endcap and args.workload != "endcap":
        line["endcap"] = endcap_section(args)
        log("endcap done")
    if args.ragged_tracks > 0 and args.workload != "ragged":
        line["ragged"] = ragged_section(args)
        log("ragged done")

    if args.serving_seconds > 0:
        if args.workload == "solenoid" and args.n_tracks is None:
            pool = events
            sdet, sfield = det, field
        else:
            sdet, sfield, pool = A.workload("solenoid", n_tracks=500_000, seed=args.seed + 1)
        line["serving"] = serving(args, sdet, sfield, pool)
        log("serving done")

    if not args.no_cpu_baseline:
        threads = max(1, min(CPU_THREADS_PER_GPU, host_cpu()["affinity_cpus"]))
        log(f"CPU baseline on {threads} threads")
        line["cpu_baseline"], ref = cpu_baseline(det, field, wl_opts, events, threads)
        log("CPU baseline done")
        got = type("GpuFit", (), {})()
        for name in ("status", "n_states", "chi2", "state_surface", "state_flags", "smoothed",
                     "smoothed_cov"):
            setattr(got, name, out64.host(name))
        got.state_flags = got.state_flags.view(np.uint32)
        line["parity_vs_cpu_oracle"] = quality.parity_counts(got, ref, events.seq_offsets)
        del ref, got
    line["bench_wall_s"] = time.perf_counter() - t_start
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
