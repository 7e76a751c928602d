"""The autotuner on the device: one fit at a time and with fits in flight
over several contexts (acts_amd.autotune.autotune(..., in_flight=k)), on a
small endcap batch.  Both return a setting of every knob from the search
space and a positive time for each candidate tried, and the tuned setting
fits the batch with the same results as the library's defaults (the knobs
change only the speed)."""
import dataclasses

import numpy as np
import pytest

import acts_amd as A
from acts_amd.autotune import SPACE, autotune

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("in_flight", [1, 2])
def test_autotune_returns_a_setting_that_fits_the_same(in_flight):
    det, field, ev = A.workload("endcap", n_tracks=3000, seed=61)
    opts = A.KalmanFitterOptions()
    print("autotune in_flight", in_flight, flush=True)
    best, ms = autotune(det, field, ev, opts, steps=1, warmup=1, cache_path="",
                        in_flight=in_flight)
    assert set(best) == set(SPACE)
    assert all(best[k] in SPACE[k] for k in SPACE)
    assert ms and all(v > 0 for v in ms.values())
    res = []
    for o in (opts, dataclasses.replace(opts, **best)):
        f = A.KalmanFitter(det, field, o)
        try:
            res.append(f.fit(ev, outputs=("state_surface", "state_flags", "smoothed")))
        finally:
            f.close()
    for name in ("status", "n_states", "chi2", "state_surface", "state_flags", "smoothed"):
        assert np.array_equal(getattr(res[0], name), getattr(res[1], name), equal_nan=True), name
