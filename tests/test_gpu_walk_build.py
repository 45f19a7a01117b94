"""GPU group walk and octree build against the CPU oracle: lists must be
identical entry for entry; the tree identical up to node numbering."""
import numpy as np
import pytest

from pyoracle import canonical_ids

pytestmark = pytest.mark.gpu

LIST_KEYS = ("groups", "cell_list", "body_list", "bodies", "cells", "origin")


def _same(a, b):
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                                                 np.ascontiguousarray(b).view(np.uint8))


@pytest.mark.parametrize("theta", [0.5, 0.7, 1.0])
@pytest.mark.parametrize("pack", [0, 1])
@pytest.mark.parametrize("gs", [1, 7, 32, 64])
def test_walk_on_host_tree_gives_identical_lists(hip, oracle, gs, pack, theta):
    pos, mass = oracle.test_system(0, 3000, 11)
    oracle.build(pos, mass, 8)
    oracle.lists(theta, gs, pack)
    want = oracle.lists_get()
    geom, links = oracle.nodes()
    tpos, tmass = oracle.tree_positions()
    hip.walk_tree(geom, links, tpos, tmass, evaluate=False, theta=theta, group_size=gs, pack=pack)
    got = hip.get_lists()
    for k in LIST_KEYS:
        assert _same(got[k], want[k]), k
    inf, winf = hip.info(), oracle.info()
    for k in ("n_groups", "n_cell_list", "n_body_list", "cell_interactions", "body_interactions",
              "max_list", "max_group"):
        assert inf[k] == winf[k], k


def _canonical_tree(geom, links):
    canon = canonical_ids(links)
    inv = np.argsort(canon)
    g = geom[inv]
    # children as canonical ids of the first child and the count
    fc = np.where(links[:, 1] > 0, canon[np.maximum(links[:, 0], 0)], -1)
    L = np.stack([fc, links[:, 1], links[:, 2], links[:, 3]], axis=1)[inv]
    return g, L


@pytest.mark.parametrize("n,kind,seed,leaf", [(3000, 0, 11, 8), (613, 1, 3, 4), (50000, 0, 3, 8),
                                              (4000, 0, 42, 1), (4000, 0, 42, 16), (9, 0, 1, 8),
                                              (1, 0, 1, 8), (2, 0, 1, 8),
                                              (2_300_000, 0, 5, 8), (20000, 0, 7, 64),
                                              (20000, 0, 7, 100)])
def test_gpu_tree_equals_reference_tree(hip, oracle, n, kind, seed, leaf):
    pos, mass = oracle.test_system(kind, n, seed)
    oracle.build(pos, mass, leaf)
    hip.group_compute(pos, mass, theta=0.7, eps=0.01, group_size=32, leaf_size=leaf, exact=True)
    geom, links, order = hip.get_tree()
    wg, wl = oracle.nodes()
    assert hip.info()["n_nodes"] == oracle.info()["n_nodes"]
    assert hip.info()["depth"] == oracle.info()["depth"]
    assert np.array_equal(order, oracle.order())
    g1, l1 = _canonical_tree(geom, links)
    g2, l2 = _canonical_tree(wg, wl)
    assert np.array_equal(l1, l2)
    assert np.array_equal(g1.view(np.uint64), g2.view(np.uint64))


def _assert_tree_equals_oracle(hip, oracle):
    geom, links, order = hip.get_tree()
    wg, wl = oracle.nodes()
    assert hip.info()["n_nodes"] == oracle.info()["n_nodes"]
    assert hip.info()["depth"] == oracle.info()["depth"]
    assert np.array_equal(order, oracle.order())
    g1, l1 = _canonical_tree(geom, links)
    g2, l2 = _canonical_tree(wg, wl)
    assert np.array_equal(l1, l2)
    assert np.array_equal(g1.view(np.uint64), g2.view(np.uint64))


def test_tree_deeper_than_the_sorted_path_levels(hip, oracle):
    """Once a context knows its tree's depth, the path sort orders by that
    many levels plus headroom only.  A tree that then needs much deeper leaves
    (a tight cluster appears) is detected and built again with every level,
    equal to the reference's; so is the shallow tree after it."""
    n = 20000
    pos, mass = oracle.test_system(0, n, 7)
    cfg = dict(theta=0.7, eps=0.01, group_size=32, leaf_size=8, exact=True)
    for _ in range(3):
        hip.group_compute(pos, mass, **cfg)
    shallow = hip.info()["depth"]
    reruns = hip.pipeline_stats()["reruns"]
    deep = pos.copy()
    rng = np.random.default_rng(3)
    deep[:300] = deep[0] + rng.normal(size=(300, 3)) * 1e-3  # depth 13 -> 17
    oracle.build(deep, mass, 8)
    hip.group_compute(deep, mass, **cfg)
    assert hip.pipeline_stats()["reruns"] > reruns
    assert hip.info()["depth"] > shallow + 3
    _assert_tree_equals_oracle(hip, oracle)
    oracle.build(pos, mass, 8)
    for _ in range(2):
        hip.group_compute(pos, mass, **cfg)
        _assert_tree_equals_oracle(hip, oracle)


@pytest.mark.parametrize("gs,pack", [(32, 1), (32, 0), (64, 1), (7, 1)])
def test_group_compute_exact_is_bitwise(hip, oracle, gs, pack):
    pos, mass = oracle.test_system(0, 20000, 7)
    want_acc, want_pot = oracle.group_gravity(pos, mass, 1.0, 0.75, 0.025, True, gs, 8, pack)
    acc, pot = hip.group_compute(pos, mass, theta=0.75, eps=0.025, group_size=gs, pack=pack,
                                 exact=True)
    assert np.array_equal(acc.view(np.uint64), want_acc.view(np.uint64))
    assert np.array_equal(pot.view(np.uint64), want_pot.view(np.uint64))
    # lists compare canonically with the oracle's (node numbering differs)
    oracle.build(pos, mass, 8)
    oracle.lists(0.75, gs, pack)
    want = oracle.lists_get()
    got = hip.get_lists()
    assert np.array_equal(got["groups"], want["groups"])
    assert np.array_equal(got["body_list"], want["body_list"])
    _, links, _ = hip.get_tree()
    _, wl = oracle.nodes()
    assert np.array_equal(canonical_ids(links)[got["cell_list"]],
                          canonical_ids(wl)[want["cell_list"]])


def test_group_compute_fast_accuracy(hip, oracle):
    # test_metal.cpp:117: median error vs FP64 direct below 3e-3 theta^2
    pos, mass = oracle.test_system(0, 5003, 5)
    acc64, _ = oracle.direct_f64(pos, mass, 1.0, 0.01)
    for theta in (0.5, 0.8):
        acc, _ = hip.group_compute(pos, mass, theta=theta, eps=0.01, group_size=32)
        e = np.sort(np.linalg.norm(acc - acc64, axis=1) / np.linalg.norm(acc64, axis=1))
        print("theta", theta, "median", e[len(e) // 2], "p99", e[len(e) * 99 // 100])
        assert e[len(e) // 2] < 3e-3 * theta * theta


def test_kdk_exact_matches_reference_steps(hip, oracle):
    pos, mass = oracle.test_system(0, 3000, 11)
    vel = np.zeros_like(pos)
    want = oracle.kdk_group(pos, vel, mass, 10, 0.01, eps=0.01)
    cfg = dict(theta=0.75, eps=0.01, group_size=32, exact=True)
    hip.state_set(pos, vel, mass)
    hip.prime(**cfg)
    hip.steps(10, 0.01, **cfg)
    got = hip.state_get()
    for a, b in zip(got, want):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_group_size_limit(hip, oracle):
    pos, mass = oracle.test_system(0, 100, 1)
    with pytest.raises(ValueError):
        hip.group_compute(pos, mass, group_size=65)


@pytest.mark.parametrize("lanes", [0, 1, 2, 3, 4, 16, 32, 64])
@pytest.mark.parametrize("gs,pack,theta", [(32, 1, 0.75), (7, 0, 0.5), (64, 1, 1.0)])
def test_walk_variants_give_identical_lists(hip, oracle, lanes, gs, pack, theta):
    # one independent walk per lane, or 16/32/64 groups sharing one traversal
    pos, mass = oracle.test_system(0, 20000, 9)
    oracle.build(pos, mass, 8)
    oracle.lists(theta, gs, pack)
    want = oracle.lists_get()
    geom, links = oracle.nodes()
    tpos, tmass = oracle.tree_positions()
    for _ in range(2):  # first call: exact offsets; second: per-group slots
        hip.walk_tree(geom, links, tpos, tmass, evaluate=False, theta=theta, group_size=gs,
                      pack=pack, walk_lanes=lanes)
        got = hip.get_lists()
        for k in ("groups", "cell_list", "body_list"):
            assert _same(got[k], want[k]), k
