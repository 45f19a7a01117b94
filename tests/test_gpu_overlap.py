"""The walk in ranges with each range's evaluation on a second stream,
alongside the walk of the next range: the same kernels on the same lists, so
forces and leapfrog states must equal the evaluation after the whole walk bit
for bit, for every number of ranges, with and without graphs, for the exact,
split-list, hi/lo and learned-slot modes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def same(a, b):
    return all(np.array_equal(bits(x), bits(y)) for x, y in zip(a, b))


MODES = {
    "exact": dict(exact=True),
    "split": dict(split_len=512),
    "hilo": dict(hilo=1),
    "tree-order": dict(eval_order=1, split_len=-1),
}


@pytest.mark.parametrize("mode", sorted(MODES))
def test_overlapped_forces_equal(hip, oracle, mode):
    pos, mass = oracle.test_system(1, 30000, 31)
    cfg = dict(theta=0.6, eps=0.01, group_size=32, **MODES[mode])
    want = hip.group_compute(pos, mass, overlap=1, **cfg)
    for ov in (2, 3, 8, 0):
        got = hip.group_compute(pos, mass, overlap=ov, **cfg)
        assert same(got, want), ov


@pytest.mark.parametrize("ov", [2, 4, 8])
def test_overlapped_graph_steps_equal(hip, oracle, ov):
    pos, mass = oracle.test_system(0, 20000, 32)
    rng = np.random.default_rng(6)
    vel = 0.3 * rng.standard_normal(pos.shape)
    cfg = dict(theta=0.7, eps=0.01, group_size=32, split_len=256)

    def run(overlap, use_graph, plan):
        hip.state_set(pos, vel, mass)
        hip.prime(overlap=overlap, **cfg)
        for k in plan:
            hip.steps(k, 0.01, use_graph=use_graph, overlap=overlap, **cfg)
        return hip.state_get()

    want = run(1, False, [5])
    before = hip.pipeline_stats()
    got = run(ov, True, [2, 3])
    after = hip.pipeline_stats()
    assert same(got, want)
    assert after["graph_steps"] - before["graph_steps"] >= 4
    t = hip.timings()
    assert t["evaluate"] > 0 and t["total"] >= t["evaluate"]
    assert same(run(ov, False, [5]), want)


def test_overlapped_predicted_slots(hip, oracle):
    # learned slots: the walk writes at predicted offsets; an overflowing range
    # is discarded and the evaluation redone with exact offsets
    pos, mass = oracle.test_system(1, 40000, 33)
    cfg = dict(theta=0.6, eps=0.01, group_size=32)
    want = hip.group_compute(pos, mass, overlap=1, **cfg)
    hip.predictor_reset(4)  # learn from the first 4 evaluations
    for _ in range(12):
        got = hip.group_compute(pos, mass, overlap=4, predict=1, **cfg)
        assert same(got, want)
    st = hip.predictor_stats()
    assert st["trained"] == 1 and st["predicted"] > 0
