"""The autotuner's search over the frame pipeline's knobs, on synthetic costs
(no GPU): coordinate descent one knob at a time, then the drain over the whole
pass-1 queue, which the descent cannot reach (tail_passes 0 needs a drain)."""
import importlib

at = importlib.import_module("chaos-ultra_amd.autotune")


def cost_model(best_point):
    """A cost that is smallest at `best_point` and grows with every knob off it."""
    def measure(p):
        return 10.0 + sum(abs(p[k] - v) for k, v in best_point.items())
    return measure


def test_descent_reaches_a_separable_optimum():
    target = {"wave_shape": 1, "block_steps": 8, "handoff_lanes": 24, "tail_passes": 3,
              "drain_refill": 0}
    best, ms, table = at.search_pipeline(cost_model(target))
    assert best == target and ms == 10.0
    # every other value of every knob at least once, then the drain-only pair
    n_desc = sum(len(v) - 1 for v in at.PIPELINE_SPACE.values())
    assert len(table) >= 1 + n_desc + len(at.DRAIN_ONLY_REFILL)
    assert [c for c, _ in table[-len(at.DRAIN_ONLY_REFILL):]] == [
        dict(target, tail_passes=0, drain_refill=r) for r in at.DRAIN_ONLY_REFILL]


def test_drain_over_the_whole_queue_is_tried():
    # the drain alone wins, but with a drain the tail passes cost more, and
    # without one tail_passes 0 is not a candidate: only the extra step finds it
    def measure(p):
        if p["tail_passes"] == 0:
            assert p["drain_refill"] > 0
            return 5.0 + abs(p["drain_refill"] - 32) / 100
        return 20.0 + (5.0 if p["drain_refill"] else 0.0)
    best, ms, table = at.search_pipeline(measure)
    assert best["tail_passes"] == 0 and best["drain_refill"] == 32 and ms == 5.0
    assert [c for c, _ in table if c["tail_passes"] == 0] == [
        dict(best, drain_refill=r) for r in at.DRAIN_ONLY_REFILL]


def test_search_never_pairs_no_tail_pass_with_no_drain():
    seen = []

    def measure(p):
        seen.append(dict(p))
        return 1.0
    at.search_pipeline(measure)
    assert all(p["tail_passes"] > 0 or p["drain_refill"] > 0 for p in seen)
    assert all(p["tail_passes"] <= 7 or p["drain_refill"] == 0 for p in seen)


def test_slow_first_measurement_does_not_lose_the_start_point():
    # the GPU's clock ramps up over a process's first renders: the start point's
    # first timing is inflated, yet a worse candidate must not beat it
    calls = []

    def measure(p):
        calls.append(dict(p))
        true = 10.0 + sum(abs(p[k] - v) for k, v in at.PIPELINE_START.items())
        return true * (1.5 if len(calls) == 1 else 1.0)
    best, ms, _ = at.search_pipeline(measure)
    assert best == at.PIPELINE_START and ms == 10.0
