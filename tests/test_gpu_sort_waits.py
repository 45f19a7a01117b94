"""The radix sorts' bounded waits: a digit scan waits for the totals the
other digits' workgroups publish, and a wait that gives up adds the digit's
total up from the tile histogram instead, so the permutation is right either
way.  A host-set test setting makes every such wait give up after a few
polls (the published totals are never accepted): trees, lists and forces
must stay bit for bit the normal ones and the oracle's, and the recounts
must be reported."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("exact", [True, False])
def test_given_up_waits_recount_and_results_stay(hip, oracle, exact):
    pos, mass = oracle.test_system(1, 30000, 9)
    cfg = dict(theta=0.7, eps=0.01, group_size=32, exact=exact)
    want = hip.group_compute(pos, mass, **cfg)
    r0 = hip.pipeline_stats()["sort_recounts"]  # (cumulative on the context)
    hip.debug_sort_waits(4)
    try:
        got = hip.group_compute(pos, mass, **cfg)
        r1 = hip.pipeline_stats()["sort_recounts"]
        _, _, order = hip.get_tree()
    finally:
        hip.debug_sort_waits(0)
    print(f"recounted digit totals: {r1 - r0}", flush=True)
    assert r1 > r0
    assert np.array_equal(bits(got[0]), bits(want[0]))
    assert np.array_equal(bits(got[1]), bits(want[1]))
    oracle.build(pos, mass, 8)
    assert np.array_equal(order, oracle.order())
    if exact:
        acc, pot = oracle.group_gravity(pos, mass, 1.0, 0.7, 0.01, True, 32, 8, True)
        assert np.array_equal(bits(got[0]), bits(acc))
    # back to the normal waits: no more recounts
    hip.group_compute(pos, mass, **cfg)
    assert hip.pipeline_stats()["sort_recounts"] == r1
