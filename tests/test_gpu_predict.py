"""The learned list-length predictor (gt_group_config.predict = 1): after a
few evaluations gather samples, the MLP is trained on the GPU and sizes every
group's slots.  The lists (and so the forces) must not depend on how the
slots were sized: predicted slots, uniform slots and the plain two-pass solver
give bit-identical states, eager or from graphs; the predictor's error and
overflow rate are reported."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def run(hip, pos, vel, mass, steps, use_graph=False, **cfg):
    hip.state_set(pos, vel, mass)
    hip.prime(**cfg)
    hip.steps(steps, 0.01, use_graph=use_graph, **cfg)
    return hip.state_get()


def moving_system(oracle, n, seed):
    pos, mass = oracle.test_system(1, n, seed)
    rng = np.random.default_rng(seed)
    return pos, 0.5 * rng.standard_normal(pos.shape), mass


@pytest.mark.parametrize("exact", [True, False])
def test_predicted_slots_change_nothing(hip, oracle, exact):
    pos, vel, mass = moving_system(oracle, 30000, 71)
    cfg = dict(theta=0.6, eps=0.01, group_size=32, exact=exact)
    want = run(hip, pos, vel, mass, 12, predict=0, **cfg)
    two_pass = run(hip, pos, vel, mass, 12, predict=2, **cfg)
    hip.predictor_reset(4)  # learn from the first 4 evaluations
    got = run(hip, pos, vel, mass, 12, predict=1, **cfg)
    st = hip.predictor_stats()
    print(st)
    assert st["trained"] == 1 and st["samples"] > 0 and st["predicted"] >= 8
    assert st["mean_abs_log_error"] < 0.5
    assert st["list_entries"] <= st["slot_entries"] or st["overflows"] > 0
    for a, b, c in zip(got, want, two_pass):
        assert np.array_equal(bits(a), bits(b)) and np.array_equal(bits(c), bits(b))


def test_predicted_slots_from_graphs(hip, oracle):
    pos, vel, mass = moving_system(oracle, 20000, 72)
    cfg = dict(theta=0.7, eps=0.01, group_size=32, predict=1)
    hip.predictor_reset(3)
    eager = run(hip, pos, vel, mass, 8, **cfg)
    hip.predictor_reset(3)
    graph = run(hip, pos, vel, mass, 8, use_graph=True, **cfg)
    st = hip.predictor_stats()
    print(st)
    assert st["trained"] == 1
    for a, b in zip(graph, eager):
        assert np.array_equal(bits(a), bits(b))


def test_predictor_is_deterministic(hip, oracle):
    pos, vel, mass = moving_system(oracle, 20000, 73)
    cfg = dict(theta=0.7, eps=0.01, group_size=32, predict=1)
    stats = []
    for _ in range(2):
        hip.predictor_reset(3)
        run(hip, pos, vel, mass, 5, **cfg)
        stats.append(hip.predictor_stats())
    assert stats[0] == stats[1]
