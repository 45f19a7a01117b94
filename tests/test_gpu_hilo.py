"""Hi/lo precision mode (gt_group_config.hilo): positions as FP32 hi + lo
pairs taken relative to each group's centre.  On a wide system (two dense
clumps 1000 units apart, softening 1e-3) FP32 positions relative to the root
centre lose ~3e-5 absolute, which is a large error on inter-particle
distances of ~1e-2; the hi/lo mode must remove it (checked against FP64
direct summation, with the multipole error made small by theta = 0.3),
leave the lists unchanged, and give graph steps equal to launched ones."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def wide_system(oracle, n, seed):
    a, ma = oracle.test_system(1, n // 2, seed)
    b, mb = oracle.test_system(1, n - n // 2, seed + 1)
    a = (a - a.mean(0)) / np.abs(a).max() + np.array([-500.0, 3.0, 1.0])
    b = (b - b.mean(0)) / np.abs(b).max() + np.array([500.0, -2.0, 0.5])
    return np.concatenate([a, b]), np.concatenate([ma, mb])


def test_hilo_removes_position_rounding(hip, oracle):
    pos, mass = wide_system(oracle, 12000, 61)
    cfg = dict(theta=0.3, eps=1e-3, group_size=32)
    ref_mode, _ = hip.group_compute(pos, mass, **cfg)
    info_ref = hip.info()
    hilo, _ = hip.group_compute(pos, mass, hilo=True, **cfg)
    info_hilo = hip.info()
    for k in ("n_groups", "cell_interactions", "body_interactions", "max_list"):
        assert info_ref[k] == info_hilo[k], k
    acc64, _ = oracle.direct_f64(pos, mass, 1.0, 1e-3)
    e_ref, e_hilo = rel_err(ref_mode, acc64), rel_err(hilo, acc64)
    print("median ref", np.median(e_ref), "hilo", np.median(e_hilo), "p99 ref",
          np.percentile(e_ref, 99), "hilo", np.percentile(e_hilo, 99))
    assert np.median(e_hilo) < 0.2 * np.median(e_ref)
    assert np.percentile(e_hilo, 99) < np.percentile(e_ref, 99)


def test_hilo_agrees_on_a_compact_system(hip, oracle):
    # where FP32 relative to the root is already fine, both modes agree to
    # FP32 rounding level
    pos, mass = oracle.test_system(0, 20000, 62)
    cfg = dict(theta=0.7, eps=0.01, group_size=32)
    a, _ = hip.group_compute(pos, mass, **cfg)
    b, _ = hip.group_compute(pos, mass, hilo=True, **cfg)
    e = rel_err(b, a)
    assert np.median(e) < 1e-5 and e.max() < 1e-3


def test_hilo_graph_steps_equal_launched(hip, oracle):
    pos, mass = wide_system(oracle, 6000, 63)
    vel = np.zeros_like(pos)
    cfg = dict(theta=0.5, eps=1e-3, group_size=32, hilo=True)
    out = []
    for use_graph in (False, True):
        hip.state_set(pos, vel, mass)
        hip.prime(**cfg)
        hip.steps(3, 0.001, use_graph=use_graph, **cfg)
        out.append(hip.state_get())
    for a, b in zip(*out):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("systems", [1, 3])
def test_hilo_group_centres_the_same_from_every_walk(hip, oracle, systems):
    # The walks over child sets write the hi/lo group centres as they compute
    # the group boxes; the other walks leave them to a kernel of their own.
    # Both round each operation the same way, so every walk gives the same
    # forces bit for bit (one system, and a batch whose systems have their
    # own origins).
    pos, mass = wide_system(oracle, 9000, 64)
    cfg = dict(theta=0.5, eps=1e-3, group_size=32, hilo=True)
    offsets = None if systems == 1 else np.array([0, 2500, 6000, 9000], np.uint64)
    got = []
    for lanes in (0, 4, 2, 1):
        if offsets is None:
            got.append(hip.group_compute(pos, mass, walk_lanes=lanes, **cfg))
        else:
            got.append(hip.group_compute_batch(offsets, pos, mass, walk_lanes=lanes, **cfg))
    for a, p in got[1:]:
        assert np.array_equal(a.view(np.uint64), got[0][0].view(np.uint64))
        assert np.array_equal(p.view(np.uint64), got[0][1].view(np.uint64))
