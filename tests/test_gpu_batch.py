"""Batches of independent systems (gt_hip_group_compute_batch and the
resident state with gt_hip_state_set_systems) against each system evaluated
on its own and against the oracle: in exact mode every system's forces must
be bit-identical to the oracle's GroupTreeGravity on that system alone
(src/gravity_backends.cpp:109-137) and to gt_hip_group_compute's for it; in
fast mode within the reference's GPU tolerance of the oracle
(test_metal.cpp:115-116); the statistics must add up; the autotuner sweeps
the configuration-4 settings on the 128-galaxy batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def systems(oracle, sizes, seed):
    """Test systems of the given sizes (kind alternating uniform / clustered),
    each shifted and scaled differently, concatenated."""
    pos, mass = [], []
    for k, n in enumerate(sizes):
        if n == 0:
            p, m = np.zeros((0, 3)), np.zeros(0)
        else:
            p, m = oracle.test_system(k % 2, n, seed + k)
            p = p * (0.5 + 0.3 * k) + np.array([10.0 * k, -3.0 * k, 1.5])
        pos.append(p)
        mass.append(m)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    return offsets, np.concatenate(pos), np.concatenate(mass)


def alone(hip, offsets, pos, mass, **cfg):
    out_a, out_p, info = [], [], []
    for s in range(len(offsets) - 1):
        b, e = int(offsets[s]), int(offsets[s + 1])
        if b == e:
            continue
        a, p = hip.group_compute(pos[b:e], mass[b:e], **cfg)
        out_a.append(a)
        out_p.append(p)
        info.append(hip.info())
    return np.concatenate(out_a), np.concatenate(out_p), info


@pytest.mark.parametrize("exact,pack,gs", [(True, 1, 32), (True, 0, 7), (False, 1, 64)])
def test_batch_equals_each_system_alone(hip, oracle, exact, pack, gs):
    # tiny systems (1, 3 and leaf + 1 bodies), an empty one, and larger ones
    sizes = [5003, 1, 0, 9, 3, 20000, 1200]
    offsets, pos, mass = systems(oracle, sizes, 31)
    cfg = dict(theta=0.7, eps=0.01, group_size=gs, pack=pack, exact=exact)
    want_a, want_p, infos = alone(hip, offsets, pos, mass, **cfg)
    got_a, got_p = hip.group_compute_batch(offsets, pos, mass, **cfg)
    inf = hip.info()
    for k in ("n_groups", "cell_interactions", "body_interactions", "n_nodes"):
        assert inf[k] == sum(i[k] for i in infos), k
    if exact:
        assert np.array_equal(bits(got_a), bits(want_a))
        assert np.array_equal(bits(got_p), bits(want_p))
    else:
        err = np.linalg.norm(got_a - want_a, axis=1) / np.maximum(
            np.linalg.norm(want_a, axis=1), 1e-30)
        assert np.median(err) < 1e-5 and err.max() < 1e-3


def test_batch_of_galaxies(hip):
    # config-4 style: galaxies of log-uniform size with scaled components
    import gt_ics
    rng = np.random.default_rng(5)
    sizes, pos, mass = [], [], []
    for k in range(12):
        n = int(np.exp(rng.uniform(np.log(2000), np.log(12000))))
        sc = rng.uniform(0.7, 1.3, 3)
        m = gt_ics.model(disk_n=int(0.4 * n), bulge_n=int(0.1 * n), halo_n=n - int(0.4 * n)
                         - int(0.1 * n), bulge_mass=0.333, bulge_rmax=2.0, disk_h=sc[0],
                         bulge_a=0.2 * sc[1], seed=100 + k)
        p, _, w, _ = gt_ics.galaxy(m)
        sizes.append(len(w))
        pos.append(p + 40.0 * k)
        mass.append(w)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    pos, mass = np.concatenate(pos), np.concatenate(mass)
    cfg = dict(theta=0.75, eps=0.025, group_size=32, pack=1, exact=True)
    want_a, want_p, _ = alone(hip, offsets, pos, mass, **cfg)
    got_a, got_p = hip.group_compute_batch(offsets, pos, mass, **cfg)
    assert np.array_equal(bits(got_a), bits(want_a))
    assert np.array_equal(bits(got_p), bits(want_p))


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_resident_steps(hip, oracle, use_graph):
    sizes = [3000, 7, 4000]
    offsets, pos, mass = systems(oracle, sizes, 41)
    rng = np.random.default_rng(2)
    vel = 0.2 * rng.standard_normal(pos.shape)
    cfg = dict(theta=0.7, eps=0.01, group_size=32, exact=True)
    want = []
    for s in range(len(sizes)):
        b, e = int(offsets[s]), int(offsets[s + 1])
        hip.state_set(pos[b:e], vel[b:e], mass[b:e])
        hip.prime(**cfg)
        hip.steps(4, 0.01, **cfg)
        want.append(hip.state_get())
    hip.state_set(pos, vel, mass)
    hip.state_set_systems(offsets)
    hip.prime(**cfg)
    hip.steps(4, 0.01, use_graph=use_graph, **cfg)
    got = hip.state_get()
    for k in range(4):
        assert np.array_equal(bits(got[k]), bits(np.concatenate([w[k] for w in want])))


def test_batch_argument_checks(hip, oracle):
    pos, mass = oracle.test_system(0, 100, 1)
    with pytest.raises(ValueError):
        hip.group_compute_batch(np.array([0, 60, 50, 100]), pos, mass)  # decreasing
    with pytest.raises(ValueError):
        hip.group_compute_batch(np.array([1, 100]), pos, mass)  # not from 0


def test_batch_128_galaxies_full_size(hip):
    # The 128-galaxy batch at its real size (about 3M particles): the fast
    # evaluation of the batch equals each galaxy evaluated alone bit for bit
    # (same trees, lists and per-group arithmetic), and every galaxy's lists
    # are counted.
    import gt_ics
    offsets, pos, _, mass = gt_ics.batch_of_galaxies()
    assert len(offsets) == 129
    cfg = dict(theta=0.75, eps=0.025, group_size=32, pack=1)
    got_a, got_p = hip.group_compute_batch(offsets, pos, mass, **cfg)
    inf = hip.info()
    want_a, want_p, infos = alone(hip, offsets, pos, mass, **cfg)
    assert inf["n_groups"] == sum(i["n_groups"] for i in infos)
    assert inf["cell_interactions"] == sum(i["cell_interactions"] for i in infos)
    assert np.array_equal(bits(got_a), bits(want_a))
    assert np.array_equal(bits(got_p), bits(want_p))


def oracle_each(oracle, offsets, pos, mass, theta, eps, gs, pack):
    acc, pot = np.zeros_like(pos), np.zeros(len(mass))
    for s in range(len(offsets) - 1):
        b, e = int(offsets[s]), int(offsets[s + 1])
        if b == e:
            continue
        acc[b:e], pot[b:e] = oracle.group_gravity(pos[b:e], mass[b:e], 1.0, theta, eps, True,
                                                  gs, 8, bool(pack))
    return acc, pot


@pytest.mark.parametrize("pack,gs", [(1, 32), (0, 7), (1, 64)])
def test_batch_exact_equals_the_oracle(hip, oracle, pack, gs):
    sizes = [5003, 1, 0, 9, 3, 20000, 1200]
    offsets, pos, mass = systems(oracle, sizes, 31)
    got_a, got_p = hip.group_compute_batch(offsets, pos, mass, theta=0.7, eps=0.01,
                                           group_size=gs, pack=pack, exact=True)
    want_a, want_p = oracle_each(oracle, offsets, pos, mass, 0.7, 0.01, gs, pack)
    assert np.array_equal(bits(got_a), bits(want_a))
    assert np.array_equal(bits(got_p), bits(want_p))


def test_batch_128_galaxies_against_the_oracle(hip, oracle):
    """Configuration 4's batch (128 galaxies, 2k-100k particles each): exact
    batched evaluation equals the oracle galaxy by galaxy bit for bit; the fast
    evaluation is within the reference's GPU tolerance of it."""
    import gt_ics
    offsets, pos, _, mass = gt_ics.batch_of_galaxies()
    cfg = dict(theta=0.75, eps=0.025, group_size=32, pack=1)
    want_a, want_p = oracle_each(oracle, offsets, pos, mass, 0.75, 0.025, 32, 1)
    got_a, got_p = hip.group_compute_batch(offsets, pos, mass, exact=True, **cfg)
    assert np.array_equal(bits(got_a), bits(want_a))
    assert np.array_equal(bits(got_p), bits(want_p))
    got_a, _ = hip.group_compute_batch(offsets, pos, mass, **cfg)
    err = np.linalg.norm(got_a - want_a, axis=1) / np.linalg.norm(want_a, axis=1)
    print(f"fast batch vs oracle: median {np.median(err):.2e} max {err.max():.2e}", flush=True)
    assert np.median(err) < 1e-5 and err.max() < 1e-3


def test_autotune_sweep_on_the_128_galaxy_batch(hip):
    """The configuration-4 sweep: group size {32, 64}, waves per SIMD, walk
    lanes, list chunk (split) length, evaluation order and the persistent grid,
    timed on the whole batch.  The settings it picks give the same forces as
    the defaults for the same group size and unsplit lists (bit for bit: tuning
    settings change how, not what, is computed), and within the fast
    evaluation's tolerance otherwise (parts of split lists are added in a fixed
    order; group size 64 makes other lists)."""
    import gt_ics
    offsets, pos, _, mass = gt_ics.batch_of_galaxies()
    cfg = dict(theta=0.75, eps=0.025, group_size=32, pack=1)
    t0 = hip.group_compute_batch(offsets, pos, mass, **cfg)
    default_ms = hip.timings()["total"]
    best, ms = hip.autotune_batch(offsets, pos, mass, search_group_size=True, **cfg)
    print(f"batch of 128: default {default_ms:.2f} ms, tuned {ms:.2f} ms with {best}", flush=True)
    assert best["group_size"] in (32, 64)
    assert best["eval_waves"] in (0, 2, 3) and best["persistent"] in (-1, 0, 1)
    assert 0 < ms <= 1.05 * default_ms
    a, p = hip.group_compute_batch(offsets, pos, mass, **dict(cfg, **best))
    split = best["split_len"]
    unsplit = split < 0 or hip.info()["max_list"] <= (split or 8192)
    if best["group_size"] == 32 and unsplit:
        assert np.array_equal(bits(a), bits(t0[0])) and np.array_equal(bits(p), bits(t0[1]))
    err = np.linalg.norm(a - t0[0], axis=1) / np.linalg.norm(t0[0], axis=1)
    assert np.median(err) < (1e-5 if best["group_size"] == 32 else 1e-3)
