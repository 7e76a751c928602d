#!/usr/bin/env python3
"""Throughput of the batched Kalman track fit on one MI355X.

One step = one fit (forward filter + RTS smoother + track summary) of the
whole synthetic batch, inputs already resident in HBM.  The timed steps
alternate between --streams fitting contexts (default 3), each on its own
stream with its own outputs, so that several fits of the batch are in flight
and one's bulk overlaps the other's last long tracks, as in a production
loop; `single_stream` gives the one-fit-at-a-time figure and the kernel
timings are per launch on one context.  Prints one JSON line.

Workloads:
  endcap (default): 100,000 muons, 10 silicon barrel cylinders (r = 32..1000
      mm, half-length 3 m) plus 2x7 endcap discs (|z| = 600..2900 mm), 0.3 mm
      Si per surface, uniform Bz = 2 T, pT 1-10 GeV, |eta| < 2.5, 15/50 um
      measurements.
  solenoid: 500,000 muons, pT log-uniform 0.5-100 GeV, same detector in a
      201x201x401 field map of a 2 T solenoid.
`value` is the plain FP64 fit (the drop-in); the mixed-precision fit (FP32
stepping and field) of the same batch is timed beside it, with its error
against the FP64 results.
"""
from __future__ import annotations

import argparse
import dataclasses
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
OUTPUTS = ("state_surface", "state_flags", "state_path", "state_chi2", "smoothed",
           "smoothed_cov")


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--streams", type=int, default=3,
                    help="fitting contexts, each on its own stream, that the timed "
                         "steps alternate between (1 = one fit at a time)")
    ap.add_argument("--workload", default="endcap",
                    choices=["endcap", "barrel", "solenoid", "ragged"])
    ap.add_argument("--n-tracks", type=int, default=None)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--input-order", action="store_true",
                    help="queue tracks in input order instead of longest first")
    ap.add_argument("--serving-seconds", type=float, default=3.0,
                    help="length of the event-serving run (0 = skip)")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the scheduling knobs on the workload before timing")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run a plain device copy of this size (for the "
                         "HBM-counter calibration of rocprof runs)")
    return ap.parse_args()


def cpu_baseline(det, field, opts, events, threads):
    """Time the CPU oracle (restatement of the reference fitter) on the host,
    on a bounded sample of the batch."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle  # noqa: E402  (test infrastructure: the timed CPU baseline)
    oracle.load()
    n = events.n_tracks
    oracle.fit(det, field, opts, events.subset(np.arange(min(n, 500))), n_threads=threads)
    ns = min(n, 100_000)
    _, secs = oracle.fit(det, field, opts, events.subset(np.arange(ns)), n_threads=threads)
    n1 = min(n, 5000)
    _, secs1 = oracle.fit(det, field, opts, events.subset(np.arange(n1)), n_threads=1)
    return {"value": ns / secs, "unit": "tracks/s", "cores": threads, "kind": "port",
            "sample": f"first {ns} tracks of the workload at {threads} threads "
                      f"(restated reference CPU fitter, g++ -O2, FP64)",
            "single_thread_value": n1 / secs1,
            "single_thread_sample": f"first {n1} tracks, 1 thread"}


def serving(args):
    """Event serving (tracks as in the field-map workload): Poisson arrivals,
    lognormal multiplicity (median 5000, clipped 500-20000).  A short
    saturated run measures the loop's capacity; events are then offered at
    half of it, for FP64 and for mixed precision."""
    import acts_amd as A
    from acts_amd import capi
    from acts_amd.serving import ServingConfig, serve
    det, field, pool = A.workload("solenoid", n_tracks=200_000, seed=args.seed + 1)
    mean_tracks = 5000 * np.exp(0.7 ** 2 / 2)
    sat = serve(det, field, pool, ServingConfig(rate_hz=1e5, duration_s=0.05, seed=args.seed),
                A.KalmanFitterOptions())
    capacity = sat["throughput_tracks_per_s"]
    rate = 0.5 * capacity / mean_tracks
    cfg = ServingConfig(rate_hz=rate, duration_s=args.serving_seconds, seed=args.seed)
    out = {"config": {"capacity_tracks_per_s": capacity, "load": 0.5,
                      "rate_hz": rate, "duration_s": cfg.duration_s,
                      "multiplicity": "lognormal median 5000 sigma 0.7 clipped 500-20000",
                      "slots": cfg.slots, "hbm_budget_bytes_per_batch": cfg.hbm_budget_bytes,
                      "tracks": "pT 0.5-100 GeV log-uniform, |eta|<2.5, 201x201x401 "
                                "solenoid map", "latency": "arrival -> results on host"}}
    for name, prec in (("fp64", capi.PRECISION_FP64), ("mixed", capi.PRECISION_MIXED)):
        out[name] = serve(det, field, pool, cfg, A.KalmanFitterOptions(precision=prec))
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
    # kernel durations of the forward filter, measured on the library's
    # stream with HIP events (one extra timed fit per sample)
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


HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
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
    from acts_amd.device import DeviceEvents, DeviceResult
    from acts_amd import quality
    from acts_amd.flops import batch_flops, forward_flops

    det, field, events = A.workload(args.workload, n_tracks=args.n_tracks, seed=args.seed)
    dev = torch.device("cuda:0")
    dev_in = DeviceEvents(events, dev)
    bin_ = dev_in.batch_in()
    stream = torch.cuda.Stream(device=dev)  # default priority
    n = events.n_tracks

    wl_opts = A.workload_options(args.workload)
    tuned = {}
    if args.autotune:
        from acts_amd.autotune import autotune
        tuned, _ = autotune(det, field, events, wl_opts, cache_path="")
    elif args.input_order:
        tuned = {"length_sorted": False}

    single = {}

    def run(precision, extra=None, outputs=OUTPUTS):
        """ms per fit with `args.streams` fits in flight (the value), plus a
        one-at-a-time timing and the per-launch kernel times of context 0."""
        opts = dataclasses.replace(wl_opts, precision=precision, kernel_timing=True, **tuned,
                                   **(extra or {}))
        outs = [DeviceResult(n, dev_in.n_slots, outputs, dev) for _ in range(args.streams)]
        fitters = [A.KalmanFitter(det, field, opts) for _ in range(args.streams)]
        ms1, wall, kt = timed_fits(fitters[0], bin_, outs[0].batch_out(), stream, args.steps,
                                   args.warmup)
        ms = ms1
        if args.streams > 1:
            ms = timed_overlapped(fitters, bin_, [o.batch_out() for o in outs], args.steps,
                                  args.warmup)
        single[(precision, tuple(sorted((extra or {}).items())))] = ms1
        for f in fitters:
            f.close()
        for o in outs[1:]:
            del o
        return ms, wall, kt, outs[0]

    ms64, wall64, kt64, out64 = run(capi.PRECISION_FP64)
    torch.cuda.synchronize()

    if args.copy_calibration_mib > 0:
        nbytes = args.copy_calibration_mib << 20
        src = torch.empty(nbytes // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()

    class R:
        pass
    r = R()
    for name in ("status", "n_steps", "n_step_trials", "n_states", "n_measurements"):
        setattr(r, name, out64.host(name))
    flop_fwd = forward_flops(r)
    flop_all = batch_flops(r)
    fwd_ms = float(np.mean([k[0] for k in kt64]))
    smooth_ms = float(np.mean([k[1] for k in kt64]))
    achieved = flop_fwd / (fwd_ms * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(REPO, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tpath) and args.n_tracks is None:
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")

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
                               f"{'field map' if args.workload == 'solenoid' else 'Bz=2T'}, "
                               "FP64 KF+RTS",
                   "n_tracks": n,
                   "fit_failures": int(np.count_nonzero(r.status)),
                   "mean_states": float(np.mean(r.n_states)),
                   "mean_steps": float(np.mean(r.n_steps))},
        "roofline": {"bound": "valu", "kernel": "fitForwardKernel (FP64)",
                     "achieved": achieved, "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_VECTOR_PEAK_TFLOPS, "traffic": traffic,
                     "flop_per_launch": flop_fwd, "kernel_ms": fwd_ms,
                     "smoother_ms": smooth_ms, "fit_flop_per_step": flop_all,
                     "forward_flop_rate_in_flight_tflops": flop_fwd / (ms64 * 1e-3) / 1e12},
        "wall_s_timed": wall64,
    }
    ms1 = single[(capi.PRECISION_FP64, ())]
    line["config"]["streams"] = args.streams
    line["single_stream"] = {"value": n / (ms1 * 1e-3), "ms_per_step": ms1,
                             "note": "one fit at a time on one stream"}

    # the Examples' configuration: fitted parameters extrapolated from the
    # first measurement to a perigee surface at the origin
    msp, _, _, outp = run(capi.PRECISION_FP64,
                          {"reference_surface": capi.REFERENCE_PERIGEE,
                           "reference_strategy": capi.EXTRAPOLATE_FIRST},
                          OUTPUTS + ("reference_params", "reference_cov"))
    line["with_perigee"] = {
        "value": n / (msp * 1e-3), "ms_per_step": msp,
        "extrapolation_failures": int(np.count_nonzero(outp.host("status")) -
                                      np.count_nonzero(r.status))}
    del outp

    line["track_states"] = track_states_bench(det, field, wl_opts, events, dev_in, stream,
                                              args.steps)

    first64 = quality.first_states(out64, events)
    # the MbfSmoother (the Examples' default smoother) in place of the RTS one
    msb, _, ktb, outb = run(capi.PRECISION_FP64, {"smoother": capi.SMOOTHER_MBF})
    line["with_mbf"] = {
        "value": n / (msb * 1e-3), "ms_per_step": msb,
        "smoother_ms": float(np.mean([k[1] for k in ktb])),
        "first_state_rel_err_vs_rts": quality.relative_error(quality.first_states(outb, events),
                                                             first64)}
    del outb
    # the SympyStepper (the Examples' stepper) in place of the EigenStepper
    mss, _, kts, outs_ = run(capi.PRECISION_FP64, {"stepper": capi.STEPPER_SYMPY})
    line["with_sympy_stepper"] = {
        "value": n / (mss * 1e-3), "ms_per_step": mss,
        "forward_kernel_ms": float(np.mean([k[0] for k in kts])),
        "first_state_rel_err_vs_eigen": quality.relative_error(
            quality.first_states(outs_, events), first64)}
    del outs_
    if not args.no_mixed:
        msx, _, ktx, outx = run(capi.PRECISION_MIXED, {"stepper": capi.STEPPER_SYMPY})
        line["with_sympy_stepper"]["mixed"] = {
            "value": n / (msx * 1e-3), "ms_per_step": msx,
            "forward_kernel_ms": float(np.mean([k[0] for k in ktx])),
            "first_state_rel_err_vs_fp64_eigen": quality.relative_error(
                quality.first_states(outx, events), first64)}
        del outx
    # the Examples' fit (KalmanFitterFunction.cpp, TrackFittingAlgorithm.cpp):
    # SympyStepper, MbfSmoother, parameters at a perigee from the first
    # measurement
    mse, _, kte, oute = run(capi.PRECISION_FP64,
                            {"stepper": capi.STEPPER_SYMPY, "smoother": capi.SMOOTHER_MBF,
                             "reference_surface": capi.REFERENCE_PERIGEE,
                             "reference_strategy": capi.EXTRAPOLATE_FIRST},
                            OUTPUTS + ("reference_params", "reference_cov"))
    line["examples_configuration"] = {
        "value": n / (mse * 1e-3), "ms_per_step": mse,
        "forward_kernel_ms": float(np.mean([k[0] for k in kte])),
        "config": "SympyStepper + MbfSmoother + perigee (first measurement)",
        "fit_failures": int(np.count_nonzero(oute.host("status")))}
    del oute
    line["pulls"] = quality.pulls(first64, events)
    if not args.no_mixed:
        msm, _, ktm, outm = run(capi.PRECISION_MIXED)
        firstm = quality.first_states(outm, events)
        line["mixed"] = {
            "value": n / (msm * 1e-3), "ms_per_step": msm, "dtype": "f32 stepping/field, f64 KF",
            "single_stream_value": n / (single[(capi.PRECISION_MIXED, ())] * 1e-3),
            "forward_kernel_ms": float(np.mean([k[0] for k in ktm])),
            "pulls": quality.pulls(firstm, events),
            "first_state_smoothed_rel_err_vs_fp64": quality.relative_error(firstm, first64)}

    if tuned:
        line["config"]["scheduling"] = tuned

    if args.serving_seconds > 0:
        line["serving"] = serving(args)

    if not args.no_cpu_baseline:
        threads = max(1, min(16, os.cpu_count() or 1))
        line["cpu_baseline"] = cpu_baseline(det, field, wl_opts, events, threads)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
