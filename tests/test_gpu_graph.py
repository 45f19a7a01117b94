"""Leapfrog steps from captured HIP graphs against the same steps launched
one by one: same kernels, same bounds, so the results must agree bit for bit
(exact and fast evaluation alike), over several calls with host work between
them, and when a step outgrows its bounds inside a graph and is finished
without graphs; prepared ahead, a call's steps run as one graph and capture
nothing.  The 50k galaxy's 10 exact steps must also reproduce the
digests of the reference's own LeapfrogKDK (tests/golden/configs.json)."""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def same_state(a, b):
    return all(np.array_equal(bits(x), bits(y)) for x, y in zip(a, b))


def run(hip, pos, vel, mass, plan, use_graph, between=None, **cfg):
    """Steps in calls of the sizes in `plan`; `between` runs after each call."""
    hip.state_set(pos, vel, mass)
    hip.prime(**cfg)
    for k in plan:
        hip.steps(k, 0.01, use_graph=use_graph, **cfg)
        if between:
            between()
    return hip.state_get()


@pytest.mark.parametrize("exact,lanes", [(True, 0), (False, 0), (True, 1), (False, 2), (False, 32)])
def test_graph_steps_equal_eager_steps(hip, oracle, exact, lanes):
    pos, mass = oracle.test_system(0, 20000, 21)
    rng = np.random.default_rng(4)
    vel = 0.3 * rng.standard_normal(pos.shape)
    cfg = dict(theta=0.7, eps=0.01, group_size=32, exact=exact, walk_lanes=lanes)
    want = run(hip, pos, vel, mass, [6], False, **cfg)
    before = hip.pipeline_stats()
    # graph launches in three calls, reading the state back in between
    got = run(hip, pos, vel, mass, [2, 1, 3], True, between=hip.state_get, **cfg)
    after = hip.pipeline_stats()
    print({k: after[k] - before[k] for k in after})
    assert same_state(got, want)
    assert after["graph_steps"] - before["graph_steps"] >= 5
    assert after["graph_captures"] > before["graph_captures"] or before["graph_captures"] > 0
    t = hip.timings()
    assert t["evaluate"] > 0 and t["total"] >= t["evaluate"]


def test_prepared_steps_run_as_one_graph(hip, oracle):
    # gt_hip_prepare_steps captures and instantiates what the next graph call
    # launches: that call then captures nothing, runs its steps as one graph
    # (events between the phases inside it), and equals eager steps bit for bit.
    pos, mass = oracle.test_system(1, 30000, 25)
    rng = np.random.default_rng(5)
    vel = 0.3 * rng.standard_normal(pos.shape)
    cfg = dict(theta=0.75, eps=0.01, group_size=32, hilo=True, persistent=1)
    want = run(hip, pos, vel, mass, [9], False, **cfg)
    hip.state_set(pos, vel, mass)
    hip.prime(**cfg)
    hip.steps(2, 0.01, use_graph=True, **cfg)
    hip.prepare_steps(7, 0.01, **cfg)
    before = hip.pipeline_stats()
    hip.steps(7, 0.01, use_graph=True, **cfg)
    after = hip.pipeline_stats()
    got = hip.state_get()
    assert same_state(got, want)
    assert after["graph_captures"] == before["graph_captures"]
    assert after["graph_steps"] - before["graph_steps"] == 7
    t = hip.timings()
    phases = t["tree"] + t["lists"] + t["evaluate"] + t["scatter"]
    print(t, flush=True)
    assert min(t["tree"], t["lists"], t["evaluate"], t["scatter"]) > 0
    assert abs(phases - t["total"]) <= 1e-3 * t["total"] + 1e-4


def test_graph_stop_and_resume(hip, oracle):
    # Slots shrunk below the previous lists: every graph step overflows its
    # slots, stops the remaining launches, and is finished without graphs
    # (lists rewritten at exact offsets).  The state must not notice.
    pos, mass = oracle.test_system(0, 8000, 22)
    vel = np.zeros_like(pos)
    cfg = dict(theta=0.75, eps=0.01, group_size=32, exact=True)
    want = run(hip, pos, vel, mass, [4], False, **cfg)
    before = hip.pipeline_stats()
    hip.debug_shrink_slots(8)
    try:
        got = run(hip, pos, vel, mass, [4], True, **cfg)
    finally:
        hip.debug_shrink_slots(0)
    after = hip.pipeline_stats()
    print({k: after[k] - before[k] for k in after})
    assert same_state(got, want)
    assert after["graph_stops"] - before["graph_stops"] >= 1
    assert after["list_rewrites"] - before["list_rewrites"] >= 1
    # and graphs work again afterwards
    got = run(hip, pos, vel, mass, [4], True, **cfg)
    assert same_state(got, want)


def test_graph_kdk_equals_golden(hip):
    import gt_ics
    with open(os.path.join(HERE, "golden", "configs.json")) as f:
        configs = json.load(f)
    c = configs["kdk50k"]
    m = gt_ics.model(bulge_mass=0.333, bulge_rmax=2.0, **configs["galaxy50k"]["model"])
    pos, _, mass, _ = gt_ics.galaxy(m)
    cfg = dict(theta=0.75, eps=0.025, group_size=32, pack=1, exact=True)
    hip.state_set(pos, np.zeros_like(pos), mass)
    hip.prime(**cfg)
    hip.steps(c["steps"], c["dt"], use_graph=True, **cfg)
    p, v, a, phi = hip.state_get()
    digest = [hashlib.sha256(np.ascontiguousarray(x).tobytes()).hexdigest() for x in (p, v, a, phi)]
    assert digest == [c["pos"], c["vel"], c["acc"], c["pot"]]


def test_bounds_grow_between_systems(hip, oracle):
    # A small uniform system, then a larger clustered one on the same context:
    # the second evaluation outgrows every bound learned from the first and is
    # run again with the sizes it reported; results equal a fresh context's.
    import gt_hip
    small, msmall = oracle.test_system(0, 3000, 23)
    big, mbig = oracle.test_system(1, 40000, 24)
    fresh = gt_hip.Context(0)
    try:
        want = fresh.group_compute(big, mbig, theta=0.6, eps=0.01, exact=True)
    finally:
        fresh.close()
    hip.group_compute(small, msmall, theta=0.6, eps=0.01, exact=True)
    got = hip.group_compute(big, mbig, theta=0.6, eps=0.01, exact=True)
    for a, b in zip(got, want):
        assert np.array_equal(bits(a), bits(b))
