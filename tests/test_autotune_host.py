"""The autotuner's search (coordinate descent over the scheduling knobs)
with a synthetic cost, and its workload signature; no device."""
import acts_amd as A
from acts_amd.autotune import DEFAULTS, SPACE, coordinate_descent, signature


def test_coordinate_descent_finds_separable_minimum():
    target = {"sched_inner_steps": SPACE["sched_inner_steps"][3],
              "sched_wait_64ths": SPACE["sched_wait_64ths"][2], "length_sorted": 3,
              "sched_quarantine_steps": SPACE["sched_quarantine_steps"][2],
              "sched_transport_skip_steps": SPACE["sched_transport_skip_steps"][1]}
    calls = []

    def cost(cfg):
        calls.append(cfg)
        c = abs(SPACE["sched_inner_steps"].index(cfg["sched_inner_steps"]) - 3)
        c += abs(SPACE["sched_wait_64ths"].index(cfg["sched_wait_64ths"]) - 2)
        c += 0 if cfg["length_sorted"] == 3 else 1
        c += abs(SPACE["sched_quarantine_steps"].index(cfg["sched_quarantine_steps"]) - 2)
        c += abs(SPACE["sched_transport_skip_steps"].index(cfg["sched_transport_skip_steps"]) - 1)
        return 1.0 + c

    best, seen = coordinate_descent(cost)
    assert best == target
    assert len(seen) == len(calls) <= 2 * sum(len(v) for v in SPACE.values()) + 1


def test_coordinate_descent_keeps_start_when_flat():
    best, _ = coordinate_descent(lambda cfg: 1.0)
    assert best == DEFAULTS


def test_signature_depends_on_shape():
    det, field, ev = A.workload("barrel", n_tracks=200, seed=2)
    s1 = signature(ev, field, A.KalmanFitterOptions())
    s2 = signature(ev.subset(range(100)), field, A.KalmanFitterOptions())
    s3 = signature(ev, field, A.KalmanFitterOptions(precision=1))
    assert len({s1, s2, s3}) == 3
