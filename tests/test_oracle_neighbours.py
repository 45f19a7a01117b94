"""The oracle's neighbour queries (Octree::for_each_within,
include/galactic_tides/octree.h:103-127), pinned: against the golden vectors
written by the reference's own code, against the reference build (when
present), against brute force (the reference's own check,
tests/test_octree.cpp:116-128: the same set as a scan of every particle), and
against an independent statement of the callback order (depth first from the
root, children last to first, tree order within a leaf)."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def same_bits(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(GOLDEN, "neighbours.npz"))


@pytest.mark.parametrize("leaf", [1, 8])
def test_within_equals_golden(oracle, golden, leaf):
    oracle.build(golden["pos"], golden["mass"], leaf)
    off, idx, d2 = oracle.within(golden["centres"], golden["radii"])
    assert same_bits(off, golden[f"leaf{leaf}_offsets"])
    assert same_bits(idx, golden[f"leaf{leaf}_index"])
    assert same_bits(d2, golden[f"leaf{leaf}_dist2"])


@pytest.mark.parametrize("kind,n,seed,leaf", [(0, 5000, 3, 8), (1, 2000, 5, 4), (0, 4000, 9, 1),
                                              (0, 4000, 9, 16)])
def test_within_equals_reference(oracle, ref, kind, n, seed, leaf):
    pos, mass = oracle.test_system(kind, n, seed)
    rng = np.random.default_rng(seed)
    centres = np.concatenate([pos[rng.integers(0, n, 200)], rng.normal(size=(50, 3))])
    for radii in (rng.uniform(0.0, 0.6, len(centres)), np.array([0.25])):
        oracle.build(pos, mass, leaf)
        ref.build(pos, mass, leaf)
        for a, b in zip(oracle.within(centres, radii), ref.within(centres, radii)):
            assert same_bits(a, b)


def test_within_brute_force(oracle):
    pos, mass = oracle.test_system(0, 3000, 21)
    rng = np.random.default_rng(4)
    centres = np.concatenate([pos[rng.integers(0, 3000, 60)], rng.normal(size=(20, 3))])
    radii = 0.05 * (1 + np.arange(len(centres)) % 10)
    oracle.build(pos, mass, 8)
    off, idx, d2 = oracle.within(centres, radii)
    for q, c in enumerate(centres):
        d = pos - c
        dd = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        want = np.nonzero(dd <= radii[q] * radii[q])[0]
        got = idx[off[q]:off[q + 1]]
        assert sorted(got.tolist()) == want.tolist()
        assert same_bits(d2[off[q]:off[q + 1]], dd[got.astype(np.int64)])


def test_within_callback_order(oracle):
    """The order, stated as a recursion: a node's children last to first, a
    leaf's particles in tree order; nodes whose cube is farther than r skipped."""
    pos, mass = oracle.test_system(1, 400, 2)
    oracle.build(pos, mass, 4)
    geom, links = oracle.nodes()
    order = oracle.order()
    tpos, _ = oracle.tree_positions()
    centres = pos[[0, 17, 99, 250]]
    radii = np.array([0.3, 1.0, 0.05, 5.0])
    off, idx, _ = oracle.within(centres, radii)
    for q, (c, r) in enumerate(zip(centres, radii)):
        seen = []

        def visit(x):
            d = np.maximum(np.abs(c - geom[x, :3]) - geom[x, 3], 0.0)
            if float(d @ d) > r * r:
                return
            if links[x, 1] == 0:
                for j in range(links[x, 2], links[x, 3]):
                    if float(np.sum((tpos[j] - c) ** 2)) <= r * r:
                        seen.append(int(order[j]))
            else:
                for k in range(links[x, 1] - 1, -1, -1):
                    visit(links[x, 0] + k)

        visit(0)
        assert seen == idx[off[q]:off[q + 1]].tolist()
        assert len(seen) > 0


def test_within_uniform_radius_is_symmetric(oracle):
    """With one radius for every query at the particles themselves, j is a
    neighbour of i exactly when i is one of j, and each particle finds itself."""
    pos, mass = oracle.test_system(0, 2000, 8)
    oracle.build(pos, mass, 8)
    off, idx, d2 = oracle.within(pos, 0.1)
    pairs = set()
    for i in range(len(pos)):
        got = idx[off[i]:off[i + 1]].tolist()
        assert i in got
        pairs.update((i, j) for j in got)
    assert all((j, i) in pairs for i, j in pairs)
