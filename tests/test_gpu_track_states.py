"""GPU parity of the track-state table (acts_kf_track_states_device, the
per-state columns of RootTrackStatesWriter) against the CPU oracle's
restatement, both fed the same device fit's outputs.

Bar: integer columns (state type, hit dimension, parameter presence,
track) identical; float columns NaN in the same places and otherwise
within 2 float ulps (the columns are doubles rounded to float; the two
math libraries' sin/cos/asinh/hypot may differ in the last double bit,
which moves a float by at most one ulp, and only where the double lies
next to a float rounding boundary).  Workloads: the endcap (uniform
field), the solenoid field map, the ragged detector with holes and
outliers, and the plane telescope with 1D strip measurements; without truth the truth-dependent columns are NaN and the
others unchanged.
"""
import numpy as np
import pytest

import acts_amd as A
from acts_amd import capi
from acts_amd import track_states as TS
from acts_amd.device import DeviceEvents, DeviceResult, to_device
from acts_amd.fitter import FitResult, TRACK_ARRAYS

pytestmark = pytest.mark.gpu

OUTPUTS = ("state_surface", "state_flags", "state_path", "state_chi2", "state_measurement",
           "predicted", "predicted_cov", "filtered", "filtered_cov", "smoothed", "smoothed_cov")


def _device_fit(workload, n, seed):
    det, field, ev = A.workload(workload, n_tracks=n, seed=seed)
    opts = A.workload_options(workload)
    fitter = A.KalmanFitter(det, field, opts)
    din = DeviceEvents(ev, "cuda:0")
    dres = DeviceResult(n, din.n_slots, OUTPUTS, "cuda:0")
    fitter.fit_device(din.batch_in(), dres.batch_out())
    fitter.sync()
    return det, field, opts, ev, fitter, din, dres


def _host_result(dres, ev):
    fr = FitResult(ev.n_tracks, int(ev.seq_offsets[-1]), OUTPUTS)
    for name in list(TRACK_ARRAYS) + list(OUTPUTS):
        arr = getattr(fr, name)
        arr[...] = dres.host(name).view(arr.dtype).reshape(arr.shape)
    return fr


def _compare(got: TS.TrackStatesTable, ref: TS.TrackStatesTable) -> dict:
    assert np.array_equal(got.ints, ref.ints)
    g, r = got.floats, ref.floats
    gn, rn = np.isnan(g), np.isnan(r)
    assert np.array_equal(gn, rn), [TS.FLOAT_COLUMNS[c] for c in np.nonzero((gn != rn).any(1))[0]]
    fin = ~rn
    ulp = np.spacing(np.maximum(np.abs(g[fin]), np.abs(r[fin])).astype(np.float32))
    d = np.abs(g[fin].astype(np.float64) - r[fin])
    assert np.all(d <= 2 * ulp), float((d / ulp).max())
    return {"values": int(fin.sum()), "not_bitwise": int(np.count_nonzero(g[fin] != r[fin]))}


@pytest.mark.parametrize("workload,n,seed", [("endcap", 4000, 61), ("solenoid", 2000, 62),
                                             ("ragged", 3000, 63), ("telescope", 3000, 65)])
def test_track_states_parity(oracle_mod, workload, n, seed):
    print("track states", workload, n, seed, flush=True)
    det, field, opts, ev, fitter, din, dres = _device_fit(workload, n, seed)
    try:
        truth = to_device(ev.truth_meas, "cuda:0")
        floats, ints = TS.device_table(fitter, din, dres, truth)
        fitter.sync()
        got = TS.host_table(floats, ints)
    finally:
        fitter.close()
    fr = _host_result(dres, ev)
    ref = oracle_mod.track_states(det, field, opts, ev, fr, ev.truth_meas)
    rep = _compare(got, ref)
    print(rep, flush=True)
    st = got.column("stateType")
    assert np.count_nonzero(st >= 0) == int(fr.n_states[fr.status == 0].sum())


def test_track_states_without_truth(oracle_mod):
    det, field, opts, ev, fitter, din, dres = _device_fit("barrel", 500, 64)
    try:
        floats, ints = TS.device_table(fitter, din, dres, None)
        fitter.sync()
        got = TS.host_table(floats, ints)
        truth = to_device(ev.truth_meas, "cuda:0")
        floats, ints = TS.device_table(fitter, din, dres, truth)
        fitter.sync()
        with_truth = TS.host_table(floats, ints)
    finally:
        fitter.close()
    ref = oracle_mod.track_states(det, field, opts, ev, _host_result(dres, ev), None)
    _compare(got, ref)
    for name in ("t_x", "res_eLOC0_smt", "pull_eQOP_ubs"):
        assert np.all(np.isnan(got.column(name)))
    for name in ("eLOC0_smt", "pull_x_hit", "err_eQOP_ubs"):
        assert np.array_equal(got.column(name), with_truth.column(name), equal_nan=True)
    assert capi.TS_NFLOAT == got.floats.shape[0]
