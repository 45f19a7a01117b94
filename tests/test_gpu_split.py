"""Long interaction lists evaluated in parts by several waves (fast
evaluation): results must stay within the reference's GPU tolerance of the
exact (CPU-identical) evaluation, be deterministic, and agree with the
unsplit evaluation to FP32 summation-order level.  Lists are made long by a
small opening angle (at theta = 0.05 nearly every cell is opened, so each
group's list approaches all N bodies: a synthetic long list)."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("theta,split", [(0.05, 512), (0.3, 256), (0.75, 64)])
def test_split_lists_match_unsplit_and_exact(hip, oracle, theta, split):
    pos, mass = oracle.test_system(0, 20000, 51)
    cfg = dict(theta=theta, eps=0.01, group_size=32)
    exact, _ = hip.group_compute(pos, mass, exact=True, **cfg)
    whole, whole_pot = hip.group_compute(pos, mass, split_len=-1, **cfg)
    inf = hip.info()
    print("theta", theta, "max list", inf["max_list"])
    parts, parts_pot = hip.group_compute(pos, mass, split_len=split, **cfg)
    again, again_pot = hip.group_compute(pos, mass, split_len=split, **cfg)
    assert inf["max_list"] > 2 * split  # the lists really were split
    # deterministic
    assert np.array_equal(bits(parts), bits(again)) and np.array_equal(bits(parts_pot),
                                                                       bits(again_pot))
    # summation order only
    e = rel_err(parts, whole)
    assert np.median(e) < 1e-6 and e.max() < 1e-4
    # the reference's GPU tolerance against the CPU kernel (test_metal.cpp:115-116)
    e = rel_err(parts, exact)
    assert np.median(e) < 1e-5 and e.max() < 1e-3


def test_nearly_direct_lists_against_fp64(hip, oracle):
    # theta = 0.05: lists of ~N entries; the grouped force is then close to
    # FP64 direct summation
    pos, mass = oracle.test_system(1, 8000, 52)
    acc, _ = hip.group_compute(pos, mass, theta=0.05, eps=0.01, group_size=64, split_len=1024)
    assert hip.info()["max_list"] > 4000
    acc64, _ = oracle.direct_f64(pos, mass, 1.0, 0.01)
    e = rel_err(acc, acc64)
    assert np.median(e) < 1e-5 and np.percentile(e, 99) < 1e-4


def test_split_steps_from_graphs_equal_launched(hip, oracle):
    pos, mass = oracle.test_system(0, 10000, 53)
    vel = np.zeros_like(pos)
    cfg = dict(theta=0.5, eps=0.01, group_size=32, split_len=128)
    out = []
    for use_graph in (False, True):
        hip.state_set(pos, vel, mass)
        hip.prime(**cfg)
        hip.steps(3, 0.01, use_graph=use_graph, **cfg)
        out.append(hip.state_get())
    for a, b in zip(*out):
        assert np.array_equal(bits(a), bits(b))


@pytest.mark.parametrize("split", [-1, 512, 128])
@pytest.mark.parametrize("persistent", [-1, 0])
def test_evaluation_order_changes_nothing(hip, oracle, split, persistent):
    # Evaluating each XCD's groups longest first only reorders whole work
    # items, and in tree order (no work items) a long list is cut into the
    # same parts by one wave, which adds them in part order as combine_parts
    # does: every target's sum is formed the same way, so bit-identical.
    pos, mass = oracle.test_system(1, 30000, 54)
    cfg = dict(theta=0.4, eps=0.01, group_size=32, split_len=split, persistent=persistent)
    got = [hip.group_compute(pos, mass, eval_order=o, **cfg) for o in (1, 2, 4, 0)]
    if split > 0:
        assert hip.info()["max_list"] > 2 * split  # the lists really were split
    for a, p in got[1:]:
        assert np.array_equal(bits(a), bits(got[0][0])) and np.array_equal(bits(p), bits(got[0][1]))


def test_autotune_picks_valid_settings(hip, oracle):
    pos, mass = oracle.test_system(1, 30000, 55)
    cfg = dict(theta=0.6, eps=0.01, group_size=32)
    best, ms = hip.autotune(pos, mass, **cfg)
    print("autotune", best, ms)
    assert ms > 0 and best["group_size"] == 32
    assert best["eval_waves"] in (0, 2, 3) and best["walk_lanes"] in (0, 2, 4, 16, 32, 64)
    assert best["eval_order"] in (0, 1, 2, 3, 4) and best["split_len"] in (0, 2048, 8192, -1)
    a0, _ = hip.group_compute(pos, mass, **cfg)
    a1, _ = hip.group_compute(pos, mass, **{**cfg, **best})
    e = rel_err(a1, a0)
    assert np.median(e) < 1e-6 and e.max() < 1e-4
    # with group sizes searched too
    best, ms = hip.autotune(pos, mass, search_group_size=True, **cfg)
    assert best["group_size"] in (32, 64)
