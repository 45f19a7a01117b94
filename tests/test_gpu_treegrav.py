"""The per-particle FP64 tree solver on the GPU (gt_hip_tree_gravity) against
the oracle's restatement of compute_gravity_tree (src/gravity.cpp:7-26,
Octree::accel_pot src/octree.cpp:275-344), itself pinned to the reference's
own build: accelerations, potentials and interaction counts bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("kind,n,seed", [(0, 5000, 41), (1, 30000, 42)])
@pytest.mark.parametrize("theta,quad,leaf", [(0.6, True, 8), (1.0, True, 16), (0.4, False, 1)])
def test_tree_gravity_equals_reference(hip, oracle, kind, n, seed, theta, quad, leaf):
    pos, mass = oracle.test_system(kind, n, seed)
    want_a, want_p, want_st = oracle.tree_gravity(pos, mass, 1.0, theta, 0.01, quad, leaf)
    got_a, got_p, got_st = hip.tree_gravity(pos, mass, G=1.0, theta=theta, eps=0.01,
                                            quadrupole=quad, leaf_size=leaf)
    assert np.array_equal(bits(got_a), bits(want_a))
    assert np.array_equal(bits(got_p), bits(want_p))
    assert np.array_equal(got_st, want_st)


def test_tree_gravity_galaxy_and_deep_clusters(hip, oracle):
    # a 50k galaxy (G = 1.3), then the same with 40 coincident particles
    # (a tree deeper than one 21-level path word)
    import gt_ics
    m = gt_ics.model(disk_n=20000, bulge_n=5000, halo_n=25000, seed=1)
    pos, _, mass, _ = gt_ics.galaxy(m)
    pos2 = pos.copy()
    pos2[:40] = pos2[100]
    for p in (pos, pos2):
        want = oracle.tree_gravity(p, mass, 1.3, 0.75, 0.025, True, 8)
        got = hip.tree_gravity(p, mass, G=1.3, theta=0.75, eps=0.025)
        for a, b in zip(got[:2], want[:2]):
            assert np.array_equal(bits(a), bits(b))
        assert np.array_equal(got[2], want[2])


def test_tree_gravity_after_group_evaluations(hip, oracle):
    # sharing a context with the grouped solver changes neither
    pos, mass = oracle.test_system(1, 20000, 43)
    g1 = hip.group_compute(pos, mass, theta=0.7, eps=0.01)
    t = hip.tree_gravity(pos, mass, theta=0.7, eps=0.01)
    g2 = hip.group_compute(pos, mass, theta=0.7, eps=0.01)
    want = oracle.tree_gravity(pos, mass, 1.0, 0.7, 0.01, True, 8)
    assert np.array_equal(bits(t[0]), bits(want[0]))
    for a, b in zip(g1, g2):
        assert np.array_equal(bits(a), bits(b))
    assert hip.tree_gravity(np.zeros((0, 3)), np.zeros(0))[2].tolist() == [0, 0]
