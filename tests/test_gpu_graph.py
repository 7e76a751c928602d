"""Fits replayed from a captured hipGraph (options.use_graph) give the eager
results bit for bit, on every replay, with eager fits of another context
and host synchronisations between the replays."""
import numpy as np
import pytest

import acts_amd as A

pytestmark = pytest.mark.gpu

NAMES = ("status", "n_states", "chi2", "n_steps", "smoothed", "smoothed_cov", "state_surface")


def test_graph_replays_match_eager():
    import torch
    from acts_amd.device import DeviceEvents, DeviceResult
    det, field, ev = A.workload("endcap", n_tracks=3000, seed=31)
    dev = torch.device("cuda:0")
    din = DeviceEvents(ev, dev)
    outs = ("state_surface", "smoothed", "smoothed_cov")
    stream = torch.cuda.Stream(device=dev)
    eager = A.KalmanFitter(det, field, A.KalmanFitterOptions())
    graph = A.KalmanFitter(det, field, A.KalmanFitterOptions(use_graph=True))
    try:
        ref = DeviceResult(ev.n_tracks, din.n_slots, outs, dev)
        eager.fit_device(din.batch_in(), ref.batch_out(), stream.cuda_stream)
        torch.cuda.synchronize()
        expect = {k: ref.host(k) for k in NAMES}
        got = DeviceResult(ev.n_tracks, din.n_slots, outs, dev)
        other = DeviceResult(ev.n_tracks, din.n_slots, outs, dev)
        for replay in range(4):
            for t in got.t.values():
                t.zero_()
            torch.cuda.synchronize()
            print("replay", replay, flush=True)
            graph.fit_device(din.batch_in(), got.batch_out(), stream.cuda_stream)
            torch.cuda.synchronize()
            for k in NAMES:
                assert np.array_equal(got.host(k), expect[k]), (replay, k)
            # eager work of another context between the replays
            eager.fit_device(din.batch_in(), other.batch_out(), stream.cuda_stream)
            torch.cuda.synchronize()
    finally:
        eager.close()
        graph.close()
