"""Host-side logic of the serving loop (no device): the Poisson/lognormal
event stream and the batch builder."""
import numpy as np

import acts_amd as A
from acts_amd.serving import ServingConfig, batch_of, event_stream


def test_event_stream_statistics():
    cfg = ServingConfig(rate_hz=2000.0, duration_s=5.0, median_tracks=5000, seed=3)
    t, m, off = event_stream(cfg, pool_tracks=100000)
    assert np.all(np.diff(t) > 0) and t[-1] < cfg.duration_s
    assert abs(len(t) / cfg.duration_s / cfg.rate_hz - 1) < 0.05
    assert m.min() >= 500 and m.max() <= 20000
    assert abs(np.median(m) / 5000 - 1) < 0.05
    assert np.all(off >= 0) and np.all(off + m <= 100000)


def test_batch_of_matches_subset():
    det, field, ev = A.workload("endcap", n_tracks=600, seed=4)
    ranges = [(10, 50), (300, 7), (0, 1), (599, 1)]
    got = batch_of(ev, ranges)
    idx = np.concatenate([np.arange(o, o + m) for o, m in ranges])
    ref = ev.subset(idx)
    for k in ("start_params", "start_cov", "start_plane", "seq_offsets", "seq_surfaces",
              "meas_offsets", "meas_surfaces", "meas_values", "meas_cov"):
        assert np.array_equal(getattr(got, k), getattr(ref, k)), k
    assert got.n_tracks == ref.n_tracks == len(idx)


def test_event_stream_fills_the_ring_in_arrival_order():
    """Each event's tracks follow the previous event's in the pool, except
    where the event would run past the end of the pool: then it starts at
    the beginning."""
    pool = 30000
    cfg = ServingConfig(rate_hz=2000.0, duration_s=1.0, median_tracks=5000, seed=9)
    t, m, off = event_stream(cfg, pool_tracks=pool)
    nxt = off[:-1] + m[:-1]
    cont = off[1:] == nxt
    wrap = off[1:] == 0
    assert np.all(cont | wrap)
    # a wrap only where the event would not have fitted after the previous one
    assert np.all(nxt[wrap & ~cont] + m[1:][wrap & ~cont] > pool)
    assert np.count_nonzero(wrap) >= 10 and np.all(off + m <= pool)
