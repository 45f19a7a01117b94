"""The CPU oracle, pinned: against the reference build (when present), against
the golden vectors written from it, against the invariants the reference's own
tests check (tests/test_gpu_kernels.cpp, tests/test_octree.cpp,
tests/test_integrator.cpp), and against closed forms."""
import hashlib
import json
import os

import numpy as np
import pytest

from pyoracle import canonical_ids

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def same_bits(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


@pytest.fixture(scope="module")
def tiny():
    return np.load(os.path.join(GOLDEN, "tiny.npz"))


@pytest.fixture(scope="module")
def configs():
    with open(os.path.join(GOLDEN, "configs.json")) as f:
        return json.load(f)


# ---- against the reference's own code ------------------------------------------------

@pytest.mark.parametrize("pack", [0, 1])
@pytest.mark.parametrize("gs", [1, 7, 32, 64])
def test_oracle_equals_reference_bitwise(oracle, ref, gs, pack):
    pos, mass = oracle.test_system(0, 3000, 11)
    for lib in (oracle, ref):
        lib.build(pos, mass, 8)
        lib.lists(0.7, gs, pack)
    g1, l1 = oracle.nodes()
    g2, l2 = ref.nodes()
    assert same_bits(g1, g2) and same_bits(l1, l2)
    assert same_bits(oracle.order(), ref.order())
    a, b = oracle.lists_get(), ref.lists_get()
    for k in a:
        assert same_bits(a[k], b[k]), k
    assert oracle.info() == ref.info()
    assert same_bits(oracle.eval(1.0, 0.01, True), ref.eval(1.0, 0.01, True))
    assert same_bits(oracle.eval(1.0, 0.01, False), ref.eval(1.0, 0.01, False))


def test_oracle_solvers_equal_reference(oracle, ref):
    pos, mass = oracle.test_system(0, 2000, 5)
    for f in ("direct_f64",):
        assert all(same_bits(x, y) for x, y in zip(getattr(oracle, f)(pos, mass, 1.0, 0.01),
                                                   getattr(ref, f)(pos, mass, 1.0, 0.01)))
    b = oracle.pack_bodies(pos, mass, np.ones(3))
    assert same_bits(b, ref.pack_bodies(pos, mass, np.ones(3)))
    assert same_bits(oracle.direct_f32(b, 1.0, 0.01), ref.direct_f32(b, 1.0, 0.01))
    t1, t2 = oracle.tree_gravity(pos, mass, 1.0, 0.6, 0.01), ref.tree_gravity(pos, mass, 1.0, 0.6,
                                                                              0.01)
    assert all(same_bits(x, y) for x, y in zip(t1, t2))
    vel = np.zeros_like(pos)
    k1 = oracle.kdk_group(pos, vel, mass, 3, 0.01, eps=0.01)
    k2 = ref.kdk_group(pos, vel, mass, 3, 0.01, eps=0.01)
    assert all(same_bits(x, y) for x, y in zip(k1, k2))
    s = np.arange(0, 2000, 11, dtype=np.uint64)
    assert same_bits(oracle.direct_f64_sample(pos, mass, s, eps=0.01),
                     ref.direct_f64_sample(pos, mass, s, eps=0.01))


def test_parallel_reference_build_is_the_same_tree(oracle, ref):
    # the reference numbers nodes differently when it builds in parallel
    # (src/octree.cpp:208-266); canonically the trees are equal
    pos, mass = oracle.test_system(0, 50000, 3)
    oracle.build(pos, mass, 8)
    ref.build(pos, mass, 8, 20000)
    g1, l1 = oracle.nodes()
    g2, l2 = ref.nodes()
    i1, i2 = np.argsort(canonical_ids(l1)), np.argsort(canonical_ids(l2))
    assert same_bits(g1[i1], g2[i2])


# ---- against the golden vectors ------------------------------------------------------

@pytest.mark.parametrize("case", ["mock307_p0", "mock307_p1", "kernels3000_p0", "kernels3000_p1",
                                  "coincident_p0", "coincident_p1"])
def test_oracle_equals_golden_tiny(oracle, tiny, case):
    pos, mass = tiny[f"{case}_pos"], tiny[f"{case}_mass"]
    leaf = int(tiny[f"{case}_leaf"][0])
    eps = float(tiny[f"{case}_eps"][0])
    theta = 0.6 if case.startswith("coincident") else 0.7
    oracle.build(pos, mass, leaf)
    oracle.lists(theta, 32, case.endswith("p1"))
    geom, links = oracle.nodes()
    assert same_bits(geom, tiny[f"{case}_geom"]) and same_bits(links, tiny[f"{case}_links"])
    assert same_bits(oracle.order(), tiny[f"{case}_order"])
    L = oracle.lists_get()
    for k in ("groups", "cell_list", "body_list"):
        assert same_bits(L[k], tiny[f"{case}_{k}"]), k
    assert same_bits(oracle.eval(1.0, eps, True), tiny[f"{case}_out"])


@pytest.mark.parametrize("name", ["galaxy50k", "galaxy200k"])
def test_oracle_equals_golden_configs(oracle, configs, name):
    import gt_ics
    c = configs[name]
    m = gt_ics.model(bulge_mass=0.333, bulge_rmax=2.0, **c["model"])
    pos, vel, mass, _ = gt_ics.galaxy(m)
    assert digest(pos) == c["positions"] and digest(mass) == c["masses"]
    oracle.build(pos, mass, 8)
    oracle.lists(0.75, 32, True)
    inf = oracle.info()
    for k in ("n_nodes", "depth", "n_groups", "n_cell_list", "n_body_list", "cell_interactions",
              "body_interactions", "max_list", "max_group"):
        assert inf[k] == c[k], k
    geom, links = oracle.nodes()
    canon = canonical_ids(links)
    L = oracle.lists_get()
    assert digest(oracle.order()) == c["order"]
    assert digest(geom[np.argsort(canon)]) == c["tree_geometry"]
    assert digest(L["groups"]) == c["groups"]
    assert digest(L["body_list"]) == c["body_list"]
    assert digest(canon[L["cell_list"]].astype(np.int64)) == c["cell_list_canonical"]
    assert digest(oracle.eval(1.0, 0.025, True)) == c["forces_fp32"]


def test_oracle_kdk_equals_golden(oracle, configs):
    import gt_ics
    c = configs["kdk50k"]
    m = gt_ics.model(bulge_mass=0.333, bulge_rmax=2.0, **configs["galaxy50k"]["model"])
    pos, _, mass, _ = gt_ics.galaxy(m)
    p, v, a, phi = oracle.kdk_group(pos, np.zeros_like(pos), mass, c["steps"], c["dt"], 1.0, 0.75,
                                    0.025, True, 32, 8, True)
    assert (digest(p), digest(v), digest(a), digest(phi)) == (c["pos"], c["vel"], c["acc"],
                                                              c["pot"])


# ---- the reference tests' invariants ----------------------------------------------------

def check_lists(links, L, n, gs):
    """test_gpu_kernels.cpp:112-133: groups tile [0, n) in order, no group
    larger than gs, every body covered exactly once per group, and no accepted
    cell contains one of its own group's targets."""
    bad = 0
    nxt = 0
    for g in L["groups"]:
        bad += int(g[0] != nxt or g[1] <= g[0] or g[1] - g[0] > gs)
        nxt = g[1]
    bad += int(nxt != n)
    for g in L["groups"]:
        cover = np.zeros(n, np.int64)
        for ci in L["cell_list"][g[2]:g[3]]:
            b, e = links[ci, 2], links[ci, 3]
            bad += int(b < g[1] and g[0] < e)
            cover[b:e] += 1
        np.add.at(cover, L["body_list"][g[4]:g[5]].astype(np.int64), 1)
        bad += int(np.count_nonzero(cover != 1))
    return bad


def check_packing(links, L):
    """test_gpu_kernels.cpp:138-168: a group is a whole node, a chunk of a leaf,
    or a run of consecutive children of the deepest node containing it."""
    bad = 0
    for g in L["groups"]:
        ni = 0
        while True:
            fc, nc = links[ni, 0], links[ni, 1]
            for c in range(nc):
                if links[fc + c, 2] <= g[0] and g[1] <= links[fc + c, 3]:
                    ni = fc + c
                    break
            else:
                break
        if (links[ni, 2], links[ni, 3]) == (g[0], g[1]) or links[ni, 1] == 0:
            continue
        kids = links[links[ni, 0]:links[ni, 0] + links[ni, 1]]
        bad += int(not ((kids[:, 2] == g[0]).any() and (kids[:, 3] == g[1]).any()))
    return bad


@pytest.mark.parametrize("pack", [False, True])
@pytest.mark.parametrize("gs", [1, 7, 32, 64])
def test_list_invariants(oracle, gs, pack):
    n = 3000
    pos, mass = oracle.test_system(0, n, 11)
    oracle.build(pos, mass, 8)
    oracle.lists(0.7, gs, pack)
    _, links = oracle.nodes()
    L = oracle.lists_get()
    inf = oracle.info()
    assert check_lists(links, L, n, gs) == 0
    assert check_packing(links, L) == 0
    mean_list = (inf["cell_interactions"] + inf["body_interactions"]) / n
    assert inf["max_list"] < 3.0 * mean_list + 64  # test_gpu_kernels.cpp:264
    assert inf["max_group"] <= gs


def test_coincident_leaf_is_chunked(oracle, tiny):
    pos, mass = tiny["coincident_p1_pos"], tiny["coincident_p1_mass"]
    oracle.build(pos, mass, 8)
    for pack in (False, True):
        oracle.lists(0.6, 32, pack)
        _, links = oracle.nodes()
        assert check_lists(links, oracle.lists_get(), len(mass), 32) == 0
        assert oracle.info()["max_group"] <= 32


def test_lists_brute_force_on_tiny_system(oracle):
    """Every node the walk reaches, brute force: for each group, a node is on
    the list iff its ancestors are all opened and it passes (or is a leaf),
    restated independently with the rule of group_lists.cpp:147-160."""
    pos, mass = oracle.test_system(1, 120, 7)
    oracle.build(pos, mass, 4)
    oracle.lists(0.7, 16, True)
    geom, links = oracle.nodes()
    tpos, _ = oracle.tree_positions()
    L = oracle.lists_get()
    theta = 0.7
    for g in L["groups"]:
        lo, hi = tpos[g[0]:g[1]].min(0), tpos[g[0]:g[1]].max(0)
        gc, gh = 0.5 * (lo + hi), 0.5 * (hi - lo)
        cells, bodies = [], []

        def visit(x):
            c, half, com = geom[x, :3], geom[x, 3], geom[x, 4:7]
            crit = 2.0 * half * (1.0 / theta) + geom[x, 14]
            overlap = bool(np.all(np.abs(c - gc) <= half * 1.0001 + gh))
            d = np.maximum(np.abs(com - gc) - gh, 0.0)
            if not overlap and float(d @ d) > crit * crit:
                if links[x, 3] - links[x, 2] <= 2:
                    bodies.extend(range(links[x, 2], links[x, 3]))
                else:
                    cells.append(x)
            elif links[x, 1] == 0:
                bodies.extend(range(links[x, 2], links[x, 3]))
            else:
                for c_ in range(links[x, 1] - 1, -1, -1):
                    visit(links[x, 0] + c_)

        visit(0)
        assert cells == list(L["cell_list"][g[2]:g[3]])
        assert bodies == list(L["body_list"][g[4]:g[5]])


def test_grouped_accuracy_not_worse_than_tree(oracle):
    # test_gpu_kernels.cpp:315-333
    n = 3000
    pos, mass = oracle.test_system(0, n, 11)
    a64, p64 = oracle.direct_f64(pos, mass, 1.0, 0.01)

    def errs(a):
        return np.sort(np.linalg.norm(a - a64, axis=1) / np.linalg.norm(a64, axis=1))

    for theta in (0.5, 0.7, 1.0):
        at, _, _ = oracle.tree_gravity(pos, mass, 1.0, theta, 0.01)
        aq, pq = oracle.group_gravity(pos, mass, 1.0, theta, 0.01, True, 32, 8, theta == 0.7)
        et, eq = errs(at), errs(aq)
        assert eq[n // 2] <= 1.05 * et[n // 2] + 2e-6
        assert eq[n * 99 // 100] <= 1.05 * et[n * 99 // 100] + 2e-5
        assert np.max(np.abs(pq / p64 - 1)) < 2e-3


# ---- closed forms ---------------------------------------------------------------------

def test_two_body_direct_closed_form(oracle):
    pos = np.array([[0.0, 0.0, 0.0], [3.0, 4.0, 0.0]])
    mass = np.array([2.0, 0.5])
    eps = 0.1
    acc, pot = oracle.direct_f64(pos, mass, 1.5, eps)
    r2 = 25.0 + eps * eps
    f0 = 1.5 * 0.5 / r2 ** 1.5 * np.array([3.0, 4.0, 0.0])
    assert np.allclose(acc[0], f0, rtol=1e-15) and np.allclose(acc[1], -f0 * 2.0 / 0.5, rtol=1e-15)
    assert np.isclose(pot[0], -1.5 * 0.5 / np.sqrt(r2), rtol=1e-15)


def test_quadrupole_closed_form(oracle):
    """Two equal bodies at +-a on x: mass 2m at the origin, traceless
    quadrupole Q_xx = 2 m (3a^2 - a^2) = 4 m a^2, Q_yy = Q_zz = -2 m a^2."""
    m, a = 0.5, 0.25
    pos = np.array([[-a, 0.0, 0.0], [a, 0.0, 0.0]])
    oracle.build(pos, np.array([m, m]), 8)
    geom, links = oracle.nodes()
    root = geom[0]
    assert root[7] == 2 * m and np.allclose(root[4:7], 0.0)
    assert np.allclose(root[8:14], [4 * m * a * a, -2 * m * a * a, -2 * m * a * a, 0, 0, 0])


def test_lone_body_feels_nothing(oracle):
    # test_gpu_kernels.cpp:237-242
    out = oracle.direct_f32(np.array([[0.5, 0.0, 0.0, 1.0]], np.float32), 1.0, 0.01)
    assert out[0, 0] == 0.0 and out[0, 3] == 0.0


def test_circular_binary_conserves_energy(oracle):
    # test_integrator.cpp:13-35: a circular binary with the group solver's
    # leapfrog keeps its energy and returns after whole orbits
    m = 0.5
    pos = np.array([[-0.5, 0.0, 0.0], [0.5, 0.0, 0.0]])
    v = np.sqrt(m / 1.0 / 2.0)  # G m / (4 r) * 2 r: circular speed at radius 0.5
    vel = np.array([[0.0, -v, 0.0], [0.0, v, 0.0]])
    mass = np.array([m, m])
    period = 2 * np.pi * 0.5 / v
    steps = 200
    p, vv, a, phi = oracle.kdk_group(pos, vel, mass, steps, period / steps, 1.0, 0.5, 0.0, True,
                                     1, 1, False)

    def energy(pp, vv_, ph):
        return 0.5 * np.sum(mass * np.sum(vv_ * vv_, 1)) + 0.5 * np.sum(mass * ph)

    _, _, _, phi0 = oracle.kdk_group(pos, vel, mass, 0, 0.0, 1.0, 0.5, 0.0, True, 1, 1, False)
    e0, e1 = energy(pos, vel, phi0), energy(p, vv, phi)
    assert abs(e1 / e0 - 1) < 1e-4
    assert np.allclose(p, pos, atol=2e-2)
