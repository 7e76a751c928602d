"""acts_kf_fit_batch_device_noalloc (the fit in the library's static
workspace, no device allocation) gives the results of
acts_kf_fit_batch_device bit for bit: on two calls in a row, the second on a
smaller batch (the workspace keeps the first call's contents), with the
length-sorted queue and the reference-surface extrapolation, and rejects a
batch that does not fit its workspace before queueing anything."""
import ctypes as C

import numpy as np
import pytest

import acts_amd as A
from acts_amd import capi
from acts_amd.fitter import ALL_STATE_OUTPUTS

pytestmark = pytest.mark.gpu

TRACK_NAMES = ("status", "n_states", "chi2", "ndf", "n_measurements", "n_holes",
               "n_outliers", "n_steps", "n_step_trials")


def _fit(fitter, din, outs, dev, stream, noalloc):
    from acts_amd.device import DeviceResult
    res = DeviceResult(din.n_tracks, din.n_slots, outs, dev)
    fn = (fitter.lib.acts_kf_fit_batch_device_noalloc if noalloc
          else fitter.lib.acts_kf_fit_batch_device)
    bin_, bout = din.batch_in(), res.batch_out()
    capi.check(fn(fitter.ctx, C.byref(bin_), C.byref(bout), stream.cuda_stream),
               "noalloc fit" if noalloc else "device fit")
    capi.check(fitter.lib.acts_kf_sync(fitter.ctx), "acts_kf_sync")
    return res


def _assert_same(got, want, names):
    for k in names:
        a, b = got.host(k), want.host(k)
        assert a.shape == b.shape, k
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), k


@pytest.mark.parametrize("with_reference", [False, True])
def test_noalloc_matches_device_entry(with_reference):
    import torch
    from acts_amd.device import DeviceEvents
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    det, field, ev = A.workload("endcap", n_tracks=3000, seed=83)
    opts = A.KalmanFitterOptions()
    outs = tuple(ALL_STATE_OUTPUTS)
    if with_reference:
        opts = A.KalmanFitterOptions(reference_surface=capi.REFERENCE_PERIGEE,
                                     reference_strategy=capi.EXTRAPOLATE_FIRST)
        outs = outs + ("reference_params", "reference_cov")
    fitter = A.KalmanFitter(det, field, opts)
    try:
        for n in (3000, 1100):  # the second call on a smaller batch
            sub = ev.subset(np.arange(n))
            din = DeviceEvents(sub, dev)
            print("noalloc: tracks", n, "reference", with_reference, flush=True)
            want = _fit(fitter, din, outs, dev, stream, noalloc=False)
            got = _fit(fitter, din, outs, dev, stream, noalloc=True)
            assert int(np.count_nonzero(want.host("status") == 0)) > 0.9 * n
            _assert_same(got, want, TRACK_NAMES + outs)
    finally:
        fitter.close()


def test_noalloc_rejects_a_batch_beyond_its_workspace():
    import torch
    det, field, ev = A.workload("endcap", n_tracks=64, seed=5)
    fitter = A.KalmanFitter(det, field, A.KalmanFitterOptions())
    try:
        dummy = torch.zeros(64, dtype=torch.float64, device="cuda:0")
        b = capi.BatchIn()
        # 10^6 tracks of 30 states need 22 GB: refused before anything runs
        b.n_tracks = 1_000_000
        b.max_sequence_length = 30
        for name in ("start_params", "start_cov", "start_plane", "seq_offsets",
                     "seq_surfaces", "meas_offsets", "meas_surfaces", "meas_values",
                     "meas_cov"):
            setattr(b, name, dummy.data_ptr())
        o = capi.BatchOut()
        rc = fitter.lib.acts_kf_fit_batch_device_noalloc(fitter.ctx, C.byref(b), C.byref(o),
                                                         None)
        assert rc == capi.ERR_INVALID_ARGUMENT
        assert "static workspace" in capi.last_error()
        capi.check(fitter.lib.acts_kf_sync(fitter.ctx), "acts_kf_sync")
    finally:
        fitter.close()
