"""The ICs generator's velocity stage with the GPU tree solver (next to the
positions, which need none): gt_hip_tree_gravity_at (Octree::accel_pot at
points, nobody left out, src/octree.cpp:275-344) equals the oracle bit for
bit, and galaxies and encounters built with the GPU's tree have the
reference generator's velocities bit for bit (make_hernquist_galaxy,
src/hernquist_ics.cpp:206-388; digests written from the reference build),
up to the 2M-particle encounter of the benchmark: the configuration inputs are
exactly what galactic-tides-ics writes."""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def configs():
    with open(os.path.join(GOLDEN, "configs.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("kind,n,theta", [(0, 5000, 0.75), (1, 20000, 0.5), (2, 3000, 1.0)])
def test_tree_gravity_at_points_equals_the_oracle(hip, oracle, kind, n, theta):
    pos, mass = oracle.test_system(kind, n, 41)
    rng = np.random.default_rng(n)
    # points inside, on and far outside the tree, and at particles themselves
    pts = np.concatenate([rng.normal(size=(2000, 3)), pos[:300], rng.normal(size=(50, 3)) * 100,
                          np.zeros((1, 3))])
    acc, pot, st = hip.tree_gravity_at(pos, mass, pts, theta=theta, eps=0.01, leaf_size=8)
    oracle.build(pos, mass, 8)
    want_acc, want_pot = oracle.accel_at(pts, 1.0, theta, 0.01, True)
    assert np.array_equal(acc.view(np.uint64), want_acc.view(np.uint64))
    assert np.array_equal(pot.view(np.uint64), want_pot.view(np.uint64))
    assert st[0] > 0


@pytest.mark.parametrize("name", ["galaxy50k", "galaxy200k", "encounter20k"])
def test_measured_velocities_with_the_gpu_tree(hip, configs, name):
    import gt_ics
    c = configs[name]
    m = gt_ics.model(bulge_mass=0.333, bulge_rmax=2.0, eps=0.025, theta=0.75, **c["model"])
    build = gt_ics.encounter if name.startswith("encounter") else gt_ics.galaxy
    pos, vel, mass, _ = build(m, tree=gt_ics.gpu_tree(hip))
    assert digest(pos) == c["positions"] and digest(mass) == c["masses"]
    assert digest(vel) == c["velocities"]


def test_the_2m_encounter_is_galactic_tides_ics_output(hip, configs):
    import gt_ics
    c = configs["encounter2m"]
    m = gt_ics.model(bulge_mass=0.333, bulge_rmax=2.0, eps=0.025, theta=0.75, **c["model"])
    pos, vel, mass, _ = gt_ics.encounter(m, tree=gt_ics.gpu_tree(hip))
    assert digest(pos) == c["positions"] and digest(mass) == c["masses"]
    assert digest(vel) == c["velocities"]
