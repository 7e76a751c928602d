"""The serving loop on the GPU: every event completes, latencies are
positive, and a track's fit does not depend on the batch it is fitted in
(alone vs batched with other events: identical results)."""
import numpy as np
import pytest

import acts_amd as A
from acts_amd.serving import SERVING_OUTPUTS, ServingConfig, batch_of, serve

pytestmark = pytest.mark.gpu


def test_serving_loop_and_batch_independence():
    det, field, pool = A.workload("endcap", n_tracks=20000, seed=41)
    cfg = ServingConfig(rate_hz=200.0, duration_s=1.0, median_tracks=2000, max_tracks=5000,
                        slots=2, seed=5)
    rep = serve(det, field, pool, cfg)
    print(rep, flush=True)
    assert rep["events"] > 100
    assert rep["latency_ms_p50"] > 0 and rep["latency_ms_p99"] >= rep["latency_ms_p50"]
    assert np.isfinite(rep["latency_ms_max"])

    fitter = A.KalmanFitter(det, field, A.KalmanFitterOptions())
    try:
        alone = fitter.fit(batch_of(pool, [(1000, 300)]), outputs=SERVING_OUTPUTS)
        mixed = fitter.fit(batch_of(pool, [(5000, 700), (1000, 300), (9000, 50)]),
                           outputs=SERVING_OUTPUTS)
    finally:
        fitter.close()
    s0 = 700  # first track of the event inside the mixed batch
    for name in ("status", "n_states", "chi2", "ndf"):
        assert np.array_equal(getattr(alone, name), getattr(mixed, name)[s0:s0 + 300]), name
    so_m = batch_of(pool, [(5000, 700)]).seq_offsets[-1]
    n_slots = alone.smoothed.shape[0]
    assert np.array_equal(alone.smoothed, mixed.smoothed[so_m:so_m + n_slots])
    assert np.array_equal(alone.state_surface, mixed.state_surface[so_m:so_m + n_slots])
