"""The persistent evaluation grid (the default, `persistent` 0 or 1: waves take
the work items of their XCD's share from a counter) runs the same per-item
code on the same items as one wave per item (`persistent = -1`), so forces
and leapfrog states must be identical
bit for bit, for split lists, hi/lo positions, tree order, and steps from
captured graphs (the counters are reset inside each replay); and the
persistent grid's forces are within the reference's GPU tolerance of the
oracle's evaluate_group_lists_cpu (test_metal.cpp:115-116)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def same(a, b):
    return all(np.array_equal(bits(x), bits(y)) for x, y in zip(a, b))


@pytest.mark.parametrize("extra", [dict(), dict(split_len=256), dict(hilo=1),
                                   dict(eval_order=1, split_len=-1), dict(eval_waves=2)])
def test_persistent_forces_equal(hip, oracle, extra):
    pos, mass = oracle.test_system(0, 40000, 17)
    cfg = dict(theta=0.6, eps=0.01, group_size=32, **extra)
    want = hip.group_compute(pos, mass, persistent=-1, **cfg)
    got = hip.group_compute(pos, mass, persistent=1, **cfg)
    default = hip.group_compute(pos, mass, **cfg)
    assert same(got, want) and same(default, want)


def test_persistent_graph_steps_equal(hip, oracle):
    pos, mass = oracle.test_system(1, 20000, 4)
    rng = np.random.default_rng(2)
    vel = 0.3 * rng.standard_normal(pos.shape)
    cfg = dict(theta=0.7, eps=0.01, group_size=32)

    def run(persistent, use_graph):
        hip.state_set(pos, vel, mass)
        hip.prime(persistent=persistent, **cfg)
        hip.steps(6, 0.01, use_graph=use_graph, persistent=persistent, **cfg)
        return hip.state_get()

    want = run(-1, False)
    before = hip.pipeline_stats()
    got = run(1, True)
    after = hip.pipeline_stats()
    assert same(got, want)
    assert after["graph_steps"] - before["graph_steps"] >= 4


@pytest.mark.parametrize("extra", [dict(), dict(split_len=256), dict(hilo=1)])
def test_persistent_against_the_oracle(hip, oracle, extra):
    pos, mass = oracle.test_system(0, 40000, 17)
    want, _ = oracle.group_gravity(pos, mass, 1.0, 0.6, 0.01, True, 32, 8, True)
    got, _ = hip.group_compute(pos, mass, theta=0.6, eps=0.01, group_size=32, persistent=1,
                               **extra)
    err = np.linalg.norm(got - want, axis=1) / np.linalg.norm(want, axis=1)
    print(f"{extra}: median {np.median(err):.2e} max {err.max():.2e}", flush=True)
    assert np.median(err) < 1e-5 and err.max() < 1e-3
