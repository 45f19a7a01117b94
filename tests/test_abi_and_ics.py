"""CPU checks of the boundary: the C-ABI library loads and exports every entry
point include/gt_hip.h declares, its structs have the reference's layouts, it
refuses cleanly without a GPU; and the ICs generator reproduces the
reference's particle positions."""
import ctypes as C
import hashlib
import json
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gt_hip.h")).read()
    return sorted(set(re.findall(r"\b(gt_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import gt_hip
    lib = C.CDLL(gt_hip.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_the_reference():
    import gt_hip
    # gravity_kernels.h:51-60 GTParams 32 B; metal_backend.mm:21-24
    assert C.sizeof(gt_hip.Params) == 32
    assert C.sizeof(gt_hip.ListsInfo) == 80
    assert C.sizeof(gt_hip.GroupConfig) == 3 * 8 + 13 * 4 + 4  # 13 ints, padded to 8


def test_compiled_in_and_clean_refusal_without_a_device():
    import gt_hip
    assert gt_hip.compiled_in()
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present here: the refusal is tested where none is")
    with pytest.raises(RuntimeError, match="no GPU"):
        gt_hip.Context(0)


def test_host_library_exports():
    import gt_ics
    for n in ("gt_ics_default_model", "gt_ics_galaxy", "gt_ics_encounter",
              "gt_ics_bench_sample"):
        assert hasattr(gt_ics._lib, n)


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", ["galaxy50k", "galaxy200k", "encounter20k"])
def test_ics_positions_equal_the_reference(name):
    import gt_ics
    with open(os.path.join(GOLDEN, "configs.json")) as f:
        c = json.load(f)[name]
    m = gt_ics.model(bulge_mass=0.333, bulge_rmax=2.0, **c["model"])
    if name.startswith("encounter"):
        pos, vel, mass, typ = gt_ics.encounter(m, 0.0, 0.0, 60.0, 30.0, 2.0, 30.0)
    else:
        pos, vel, mass, typ = gt_ics.galaxy(m)
    assert _digest(pos) == c["positions"]
    assert _digest(mass) == c["masses"]
    assert np.isfinite(vel).all()


def test_bench_sample_is_the_reference_benchmarks():
    """The force-error sample (apps/galactic_tides_bench.cpp:74-78) at the 2M
    encounter equals the one the reference's code drew for the golden data."""
    import gt_ics
    g = np.load(os.path.join(GOLDEN, "encounter2m.npz"))
    sample = gt_ics.bench_sample(2_000_000, 1000, 1)
    assert sample.dtype == np.uint64 and np.array_equal(sample, g["sample"])
    assert len(np.unique(sample)) == 1000


@pytest.mark.parametrize("n,size", [(1, 1000), (10, 3), (50_000, 1000), (200_000, 1000)])
def test_bench_sample_small_sizes_against_the_reference(ref, n, size):
    import gt_ics
    assert np.array_equal(gt_ics.bench_sample(n, size, 1), ref.bench_sample(n, size, 1))


def test_ics_component_order_and_centre_of_mass():
    import gt_ics
    m = gt_ics.model(disk_n=300, bulge_n=100, halo_n=600, seed=7)
    pos, vel, mass, typ = gt_ics.galaxy(m)
    # halo, then disk, then bulge (src/hernquist_ics.cpp:175-204), GADGET types
    assert (typ[:600] == 1).all() and (typ[600:900] == 2).all() and (typ[900:] == 3).all()
    com = (mass[:, None] * pos).sum(0) / mass.sum()
    cv = (mass[:, None] * vel).sum(0) / mass.sum()
    assert np.abs(com).max() < 1e-12 and np.abs(cv).max() < 1e-12
    assert np.isclose(mass.sum(), m.disk_mass + m.bulge_mass + m.halo_mass)


def oracle_tree(chk):
    """The velocity stage's tree solver from a CPU checker (compute_gravity_tree
    and Octree::accel_pot at points, restated; pinned to the reference)."""
    def solve(m, pos, mass, points):
        if points is None:
            acc, pot, _ = chk.tree_gravity(pos, mass, m.G, m.theta, m.eps, bool(m.quadrupole), 8)
            return acc, pot
        chk.build(pos, mass, 8)
        return chk.accel_at(points, m.G, m.theta, m.eps, bool(m.quadrupole))
    return solve


@pytest.mark.parametrize("name", ["galaxy50k", "galaxy200k", "encounter20k"])
def test_measured_velocities_equal_the_reference(oracle, name):
    """The reference's velocity stage (src/hernquist_ics.cpp:206-388) with the
    oracle's tree: make_hernquist_galaxy's velocities bit for bit (digests
    written from the reference build), eps 0.025 and theta 0.75 as the
    configurations' ICs measure their potential."""
    import gt_ics
    with open(os.path.join(GOLDEN, "configs.json")) as f:
        c = json.load(f)[name]
    m = gt_ics.model(bulge_mass=0.333, bulge_rmax=2.0, eps=0.025, theta=0.75, **c["model"])
    if name.startswith("encounter"):
        pos, vel, mass, typ = gt_ics.encounter(m, tree=oracle_tree(oracle))
    else:
        pos, vel, mass, typ = gt_ics.galaxy(m, tree=oracle_tree(oracle))
    assert _digest(pos) == c["positions"] and _digest(mass) == c["masses"]
    assert _digest(vel) == c["velocities"]


def test_measured_velocities_small_against_the_reference_build(ref):
    """Every particle type's velocities against the reference's generator
    itself, and a tree callback that fails is reported."""
    import sys
    import gt_ics
    sys.path.insert(0, GOLDEN)
    from make_fixtures import ref_galaxy
    rp, rv, rm, rt = ref_galaxy(ref, 1200, 300, 1500, 9)
    m = gt_ics.model(disk_n=1200, bulge_n=300, halo_n=1500, bulge_mass=0.333, bulge_rmax=2.0,
                     seed=9, eps=0.025, theta=0.75)
    p, v, ms, t = gt_ics.galaxy(m, tree=oracle_tree(ref))
    assert np.array_equal(p, rp) and np.array_equal(v.view(np.uint64), rv.view(np.uint64))
    assert np.array_equal(t, rt)

    def broken(m, pos, mass, points):
        raise RuntimeError("tree unavailable")
    with pytest.raises(RuntimeError, match="tree unavailable"):
        gt_ics.galaxy(m, tree=broken)
