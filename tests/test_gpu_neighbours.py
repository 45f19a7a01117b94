"""GPU neighbour queries (gt_hip_neighbours, Octree::for_each_within) against
the oracle and the golden vectors: offsets, original indices and squared
distances identical, in the reference's callback order; edge cases (no
queries, no particles, one particle, zero / negative / NaN / all-covering
radii, centres outside the root cube, a too small max_out); and at 2.3M
particles, the oracle over every query plus size-independent checks."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def same_bits(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def check_same(got, want):
    for a, b, what in zip(got, want, ("offsets", "index", "dist2")):
        assert same_bits(a, b), what


@pytest.mark.parametrize("leaf", [1, 8])
def test_neighbours_equal_golden(hip, leaf):
    g = np.load(os.path.join(GOLDEN, "neighbours.npz"))
    got = hip.neighbours(g["pos"], g["centres"], g["radii"], mass=g["mass"], leaf_size=leaf)
    check_same(got, (g[f"leaf{leaf}_offsets"], g[f"leaf{leaf}_index"], g[f"leaf{leaf}_dist2"]))


@pytest.mark.parametrize("kind,n,seed,leaf", [(0, 3000, 11, 8), (1, 613, 3, 4), (0, 50000, 3, 8),
                                              (0, 4000, 42, 1), (0, 4000, 42, 16)])
def test_neighbours_equal_oracle(hip, oracle, kind, n, seed, leaf):
    pos, mass = oracle.test_system(kind, n, seed)
    rng = np.random.default_rng(seed)
    centres = np.concatenate([pos[rng.integers(0, n, 500)], rng.normal(size=(100, 3)) * 2,
                              rng.normal(size=(20, 3)) * 100])
    oracle.build(pos, mass, leaf)
    for radii in (rng.uniform(0.0, 0.5, len(centres)), np.array([0.1])):
        got = hip.neighbours(pos, centres, radii, mass=mass, leaf_size=leaf)
        check_same(got, oracle.within(centres, radii))
        assert hip.info()["n_nodes"] == oracle.info()["n_nodes"]
        assert hip.info()["depth"] == oracle.info()["depth"]


def test_neighbours_without_masses(hip, oracle):
    pos, mass = oracle.test_system(0, 5000, 1)
    oracle.build(pos, mass, 8)
    check_same(hip.neighbours(pos, pos[:300], 0.2), oracle.within(pos[:300], 0.2))


def test_neighbours_edge_radii(hip, oracle):
    pos, mass = oracle.test_system(0, 2000, 6)
    oracle.build(pos, mass, 8)
    centres = np.concatenate([pos[:3], [[0.0, 0.0, 0.0], [1e6, 0.0, 0.0]], pos[3:6]])
    radii = np.array([0.0, -0.3, np.nan, 1e9, 1e3, np.inf, 0.2, 1e-300])
    got = hip.neighbours(pos, centres, radii, mass=mass)
    check_same(got, oracle.within(centres, radii))
    off = got[0]
    assert off[1] - off[0] >= 1  # radius 0 at a particle finds it
    assert off[3] == off[2]      # NaN: nothing
    assert off[4] - off[3] == len(pos) and off[6] - off[5] == len(pos)


def test_neighbours_small_and_empty(hip, oracle):
    off, idx, d2 = hip.neighbours(np.zeros((0, 3)), np.ones((4, 3)), 1.0)
    assert off.tolist() == [0] * 5 and len(idx) == 0 and len(d2) == 0
    off, idx, d2 = hip.neighbours(np.ones((10, 3)), np.zeros((0, 3)), 1.0)
    assert off.tolist() == [0] and len(idx) == 0
    one = np.array([[0.5, -0.25, 2.0]])
    off, idx, d2 = hip.neighbours(one, np.array([[0.5, -0.25, 2.0], [0.5, 0.0, 2.0]]),
                                  np.array([0.0, 0.2]))
    assert off.tolist() == [0, 1, 1] and idx.tolist() == [0] and d2.tolist() == [0.0]
    pos = np.zeros((150, 3))  # coincident particles: one leaf deeper than any split
    pos[:100] = (1.0, 2.0, 3.0)
    i = np.arange(50, dtype=np.float64)
    pos[100:] = np.stack([i, -i, 0.5 * i], 1)
    mass = np.ones(150)
    oracle.build(pos, mass, 8)
    c = np.array([[1.0, 2.0, 3.0], [10.0, -10.0, 5.0]])
    check_same(hip.neighbours(pos, c, 1.5, mass=mass), oracle.within(c, 1.5))


def test_neighbours_max_out(hip, oracle):
    pos, mass = oracle.test_system(0, 3000, 2)
    want = hip.neighbours(pos, pos[:100], 0.3)
    total = int(want[0][-1])
    with pytest.raises(OverflowError):
        hip.neighbours(pos, pos[:100], 0.3, max_out=total - 1)
    check_same(hip.neighbours(pos, pos[:100], 0.3, max_out=total), want)
    with pytest.raises(ValueError):
        hip.neighbours(pos, pos[:100], np.array([0.1, 0.2]))  # neither 1 nor nq radii
    with pytest.raises(ValueError):
        hip.neighbours(pos, pos[:100], 0.3, leaf_size=0)


def test_neighbours_2m(hip, oracle):
    """2.3M particles, 400k queries at particles (uniform radius, about 30
    neighbours each) and 20k at random points: identical to the oracle; every
    particle query finds itself at distance 0; the relation is symmetric."""
    n = 2_300_000
    pos, mass = oracle.test_system(0, n, 5)
    rng = np.random.default_rng(12)
    sel = rng.choice(n, 400_000, replace=False)
    centres = np.concatenate([pos[sel], rng.normal(size=(20_000, 3))])
    r = 0.02
    got = hip.neighbours(pos, centres, r, mass=mass)
    oracle.build(pos, mass, 8)
    check_same(got, oracle.within(centres, r))
    off, idx, d2 = got
    counts = np.diff(off.astype(np.int64))
    print("neighbours per particle query", counts[:400_000].mean(), "total", off[-1])
    starts = off[:400_000].astype(np.int64)
    # each particle query contains the particle itself, at distance zero
    q_of = np.repeat(np.arange(len(centres)), counts)
    hit_self = (idx == sel[np.minimum(q_of, 399_999)]) & (q_of < 400_000)
    assert np.bincount(q_of[hit_self], minlength=400_000)[:400_000].min() >= 1
    assert np.all(d2[hit_self] == 0.0)
    # symmetry among the selected particles: i finds j iff j finds i
    is_sel = np.zeros(n, bool)
    is_sel[sel] = True
    m = (q_of < 400_000) & is_sel[idx]
    a = sel[q_of[m]].astype(np.int64) * n + idx[m].astype(np.int64)
    b = idx[m].astype(np.int64) * n + sel[q_of[m]].astype(np.int64)
    assert np.array_equal(np.sort(a), np.sort(b))
    assert starts[0] == 0
