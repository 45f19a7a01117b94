"""The resident leapfrog state is kept in the order of the tree of its last
priming (gt_hip_prime), so gathers into tree order are nearly sequential.
None of that may show: downloads, forces, the tree's body order and a later
change of systems all see the caller's particle order, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFG = dict(theta=0.75, eps=0.01, group_size=32, leaf_size=8, exact=True)


def same_bits(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def test_prime_and_steps_keep_the_callers_order(hip, oracle):
    n = 6000
    pos, mass = oracle.test_system(0, n, 9)
    vel = np.random.default_rng(1).normal(size=pos.shape) * 0.05
    hip.state_set(pos, vel, mass)
    hip.prime(**CFG)
    p, v, a, q = hip.state_get()
    assert same_bits(p, pos) and same_bits(v, vel)
    acc, pot = hip.group_compute(pos, mass, **CFG)
    assert same_bits(a, acc) and same_bits(q, pot)
    # the tree of the priming, in caller indices
    hip.state_set(pos, vel, mass)
    hip.prime(**CFG)
    _, _, order = hip.get_tree()
    oracle.build(pos, mass, 8)
    assert np.array_equal(order, oracle.order())
    # steps on the permuted state: the tree of the last step, and the state,
    # against the same steps on an unpermuted start
    hip.steps(3, 0.01, use_graph=True, **CFG)
    p3, v3, a3, q3 = hip.state_get()
    _, _, order = hip.get_tree()
    want = oracle.kdk_group(pos, vel, mass, 3, 0.01, eps=0.01)
    for got, ref in zip((p3, v3, a3, q3), want):
        assert same_bits(got, ref)
    oracle.build(p3, mass, 8)  # the last step's tree (the final kick moves nothing)
    assert np.array_equal(order, oracle.order())


def test_systems_set_after_priming(hip, oracle):
    n = 4000
    pos, mass = oracle.test_system(0, n, 4)
    vel = np.zeros_like(pos)
    hip.state_set(pos, vel, mass)
    hip.prime(**CFG)  # the state is now in this tree's order
    off = np.array([0, 1500, n])
    hip.state_set_systems(off)  # ranges of the caller's order
    hip.prime(**CFG)
    p, v, a, q = hip.state_get()
    assert same_bits(p, pos) and same_bits(v, vel)
    acc, pot = hip.group_compute_batch(off, pos, mass, **CFG)
    assert same_bits(a, acc) and same_bits(q, pot)
    d = hip.state_diagnostics(0.0, systems=2)
    assert len(d) == 2
    for s in range(2):
        m = mass[off[s]:off[s + 1]]
        com = (m[:, None] * pos[off[s]:off[s + 1]]).sum(0) / m.sum()
        assert np.allclose(d[s]["centre_of_mass"], com, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("graph", [True, False])
def test_reorders_during_long_runs(hip, oracle, graph):
    """The state is put back into tree order every 64 steps within a run: 70
    steps, graph or eager, equal the reference's steps bit for bit."""
    n = 2000
    pos, mass = oracle.test_system(0, n, 13)
    vel = np.random.default_rng(2).normal(size=pos.shape) * 0.3
    hip.state_set(pos, vel, mass)
    hip.prime(**CFG)
    hip.steps(70, 0.005, use_graph=graph, **CFG)
    got = hip.state_get()
    want = oracle.kdk_group(pos, vel, mass, 70, 0.005, eps=0.01)
    for a, b in zip(got, want):
        assert same_bits(a, b)
    _, _, order = hip.get_tree()
    oracle.build(got[0], mass, 8)
    assert np.array_equal(order, oracle.order())


def test_batch_reorders_during_long_runs(hip, oracle):
    """A batch of systems stepped 70 times from graphs (a reorder of the state
    at 64 steps in the middle of the run) equals each system stepped alone
    eagerly, and the oracle's steps, bit for bit."""
    sizes = [2500, 9, 3000]
    pos, mass, vel = [], [], []
    for k, n in enumerate(sizes):
        p, m = oracle.test_system(k % 2, n, 51 + k)
        pos.append(p + 6.0 * k)
        mass.append(m)
        vel.append(np.random.default_rng(k).normal(size=p.shape) * 0.2)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    P, M, V = np.concatenate(pos), np.concatenate(mass), np.concatenate(vel)
    hip.state_set(P, V, M)
    hip.state_set_systems(off)
    hip.prime(**CFG)
    hip.steps(70, 0.005, use_graph=True, **CFG)
    got = hip.state_get()
    for s, n in enumerate(sizes):
        b, e = int(off[s]), int(off[s + 1])
        hip.state_set(pos[s], vel[s], mass[s])
        hip.prime(**CFG)
        hip.steps(70, 0.005, use_graph=False, **CFG)
        alone = hip.state_get()
        want = oracle.kdk_group(pos[s], vel[s], mass[s], 70, 0.005, eps=0.01)
        for g, a, w in zip(got, alone, want):
            assert same_bits(g[b:e], a) and same_bits(a, w)


def test_fast_graph_steps_equal_eager_over_long_runs(hip, oracle):
    """The fast evaluation (v_rsq_f32, FMA, the wave evaluator's fixed
    summation order) stepped 70 times: from graphs (with the reorder at 64
    steps) bit for bit what launched kernels give."""
    n = 5000
    pos, mass = oracle.test_system(1, n, 21)
    vel = np.random.default_rng(3).normal(size=pos.shape) * 0.3
    cfg = dict(CFG, exact=False)
    runs = []
    for graph in (False, True):
        hip.state_set(pos, vel, mass)
        hip.prime(**cfg)
        hip.steps(70, 0.005, use_graph=graph, **cfg)
        runs.append(hip.state_get())
    for a, b in zip(*runs):
        assert same_bits(a, b)
