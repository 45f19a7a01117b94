"""Writes the golden vectors in tests/golden/ from the reference itself.

Runs only where the reference build exists (oracle/_ref/libgt_ref.so, built
by `make -C oracle ref` from /root/reference): every expected value below is
produced by the reference's own code (Octree::build, build_group_lists,
evaluate_group_lists_cpu, LeapfrogKDK with GroupTreeGravity, make_hernquist_galaxy).
The tests compare the CPU oracle and the GPU library against these files.

    python tests/golden/make_fixtures.py [--only neighbours]

Outputs:
  tiny.npz        full inputs and outputs for three small systems
  configs.json    SHA-256 digests and statistics for the 50k and 200k galaxies,
                  a small encounter, and 10 leapfrog steps on the 50k galaxy
  encounter2m.npz the 2M-particle encounter (seeds 3 and 4): the benchmark's
                  force-error sample (apps/galactic_tides_bench.cpp:74-78) and FP64
                  direct-summation accelerations at it (src/gravity.cpp:28-45), for
                  the reference's halo and for the configurations' Hernquist halo;
                  its summary (counts, digests of trees, lists and FP32 forces) goes
                  into configs.json under "encounter2m"
  neighbours.npz  Octree::for_each_within results (CSR: offsets, original
                  indices, squared distances, in callback order) on the 3000-particle
                  system of test_gpu_kernels.cpp, leaf sizes 1 and 8
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from pyoracle import Checker, canonical_ids  # noqa: E402


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def model_vector(disk_n, bulge_n, halo_n, seed, eps=0.025, theta=0.75):
    """HernquistModel fields in ref_shim.cpp's order, with the configurations'
    bulge mass 0.333 and cut-off radius 2."""
    return np.array([1.0, disk_n, 1.0, 1.0, 0.2, 10.0, 1.5, 2.43, bulge_n, 0.333, 0.2, 2.0,
                     halo_n, 5.8, 1.0, 10.0, 30.0, 0.95, seed, eps, theta], np.float64)


def ref_galaxy(ref, disk_n, bulge_n, halo_n, seed):
    n = disk_n + bulge_n + halo_n
    pos, vel, mass = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    typ = np.zeros(n, np.uint8)
    ref._galaxy(model_vector(disk_n, bulge_n, halo_n, seed), 0.0, 0.0, pos, vel, mass, typ)
    return pos, vel, mass, typ


def ref_encounter(ref, disk_n, bulge_n, halo_n, seed):
    n = 2 * (disk_n + bulge_n + halo_n)
    pos, vel, mass = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    typ = np.zeros(n, np.uint8)
    ref._encounter(model_vector(disk_n, bulge_n, halo_n, seed), 0.0, 0.0, 60.0, 30.0, 2.0, 30.0,
                   pos, vel, mass, typ)
    return pos, vel, mass, typ


def tree_and_lists(ref, pos, mass, leaf, theta, gs, pack, eps):
    ref.build(pos, mass, leaf)
    geom, links = ref.nodes()
    order = ref.order()
    ref.lists(theta, gs, pack)
    L = ref.lists_get()
    out = ref.eval(1.0, eps, True)
    return geom, links, order, L, out, ref.info()


def summary(ref, pos, mass, theta=0.75, gs=32, pack=True, eps=0.025, leaf=8):
    geom, links, order, L, out, info = tree_and_lists(ref, pos, mass, leaf, theta, gs, pack, eps)
    canon = canonical_ids(links)
    inv = np.argsort(canon)
    fc = np.where(links[:, 1] > 0, canon[np.maximum(links[:, 0], 0)], -1)
    clinks = np.stack([fc, links[:, 1], links[:, 2], links[:, 3]], 1)[inv].astype(np.int64)
    return {
        "positions": digest(pos), "masses": digest(mass),
        "n_nodes": info["n_nodes"], "depth": info["depth"], "n_groups": info["n_groups"],
        "n_cell_list": info["n_cell_list"], "n_body_list": info["n_body_list"],
        "cell_interactions": info["cell_interactions"],
        "body_interactions": info["body_interactions"],
        "max_list": info["max_list"], "max_group": info["max_group"],
        "order": digest(order.astype(np.uint64)),
        "tree_geometry": digest(geom[inv]), "tree_links": digest(clinks),
        "groups": digest(L["groups"]), "body_list": digest(L["body_list"]),
        "cell_list_canonical": digest(canon[L["cell_list"]].astype(np.int64)),
        "forces_fp32": digest(out),
    }


def neighbour_queries(pos):
    """Query centres and radii: particles as centres with the radii of the
    reference's neighbour test (0.05 (1 + k mod 10), tests/test_octree.cpp:116-128),
    seeded random centres, a zero radius and one radius that takes in every
    particle."""
    rng = np.random.default_rng(7)
    pick = rng.integers(0, len(pos), 40)
    centres = [pos[pick], rng.normal(size=(10, 3)) * 0.5, pos[[5]], np.zeros((1, 3))]
    radii = [0.05 * (1 + np.arange(40) % 10), np.full(10, 0.3), [0.0], [1e3]]
    return np.concatenate(centres), np.concatenate(radii).astype(np.float64)


def write_neighbours(ref):
    oracle = Checker("oracle")
    pos, mass = oracle.test_system(0, 3000, 11)
    centres, radii = neighbour_queries(pos)
    out = {"pos": pos, "mass": mass, "centres": centres, "radii": radii}
    for leaf in (1, 8):
        ref.build(pos, mass, leaf)
        off, idx, d2 = ref.within(centres, radii)
        out[f"leaf{leaf}_offsets"], out[f"leaf{leaf}_index"], out[f"leaf{leaf}_dist2"] = off, idx, d2
    np.savez_compressed(os.path.join(HERE, "neighbours.npz"), **out)
    print("wrote neighbours.npz")


def sample_errors(acc, ref_acc):
    """Median and p99 of |a - a64| / |a64| (apps/galactic_tides_bench.cpp:142-149)."""
    e = np.sort(np.linalg.norm(acc - ref_acc, axis=1) / np.linalg.norm(ref_acc, axis=1))
    return float(e[len(e) // 2]), float(e[len(e) * 99 // 100])


def write_encounter2m(ref):
    """Configuration 3 at full size.  Reference halo: everything from the
    reference's own code (ICs, tree, lists, FP32 evaluation, FP64 direct).
    Hernquist halo (the configurations' halo, which the reference's ICs do not
    offer): positions from this repository's generator, FP64 direct from the
    reference's code."""
    sys.path.insert(0, os.path.join(ROOT, "galactic-tides_amd"))
    import gt_ics
    pos, vel, mass, typ = ref_encounter(ref, 400000, 100000, 500000, 3)
    n = len(mass)
    c = summary(ref, pos, mass)
    c["model"] = dict(disk_n=400000, bulge_n=100000, halo_n=500000, seed=3)
    c["velocities"] = digest(vel)
    sample = ref.bench_sample(n, 1000, 1)
    a64 = ref.direct_f64_sample(pos, mass, sample, 1.0, 0.025)
    # the reference's own FP32 grouped evaluation (tree order) at the sample
    out = ref.eval(1.0, 0.025, True)
    order = ref.order().astype(np.int64)
    acc = np.empty((n, 3))
    acc[order] = out[:, :3].astype(np.float64)
    c["sample_median"], c["sample_p99"] = sample_errors(acc[sample.astype(np.int64)], a64)
    print("encounter2m", c["n_nodes"], c["n_groups"], c["n_cell_list"], c["n_body_list"],
          "p50/p99", c["sample_median"], c["sample_p99"], flush=True)
    m = gt_ics.model(disk_n=400000, bulge_n=100000, halo_n=500000, bulge_mass=0.333,
                     bulge_a=0.2, bulge_rmax=2.0, halo_mass=5.8, halo_rmax=30.0,
                     halo_profile="hernquist", halo_a=4.0, seed=3)
    hpos, _, hmass, _ = gt_ics.encounter(m, 0.0, 0.0, 60.0, 30.0, 2.0, 30.0)
    h64 = ref.direct_f64_sample(hpos, hmass, sample, 1.0, 0.025)
    hs = summary(ref, hpos, hmass)
    hout = ref.eval(1.0, 0.025, True)
    horder = ref.order().astype(np.int64)
    hacc = np.empty((n, 3))
    hacc[horder] = hout[:, :3].astype(np.float64)
    hmed, hp99 = sample_errors(hacc[sample.astype(np.int64)], h64)
    c["hernquist"] = {k: hs[k] for k in ("positions", "masses", "n_nodes", "n_groups",
                                         "cell_interactions", "body_interactions", "groups",
                                         "body_list", "cell_list_canonical", "forces_fp32")}
    c["hernquist"].update(sample_median=hmed, sample_p99=hp99)
    print("encounter2m hernquist p50/p99", hmed, hp99, flush=True)
    np.savez_compressed(os.path.join(HERE, "encounter2m.npz"), sample=sample, acc64=a64,
                        hernquist_acc64=h64)
    return c


def add_velocities(ref, configs):
    """Digests of make_hernquist_galaxy's velocities (src/hernquist_ics.cpp:206-388,
    measured with the reference's tree at eps 0.025, theta 0.75) for the
    galaxies and the small encounter."""
    for name in ("galaxy50k", "galaxy200k"):
        md = configs[name]["model"]
        pos, vel, mass, typ = ref_galaxy(ref, md["disk_n"], md["bulge_n"], md["halo_n"], md["seed"])
        assert digest(pos) == configs[name]["positions"]
        configs[name]["velocities"] = digest(vel)
    md = configs["encounter20k"]["model"]
    pos, vel, mass, typ = ref_encounter(ref, md["disk_n"], md["bulge_n"], md["halo_n"], md["seed"])
    assert digest(pos) == configs["encounter20k"]["positions"]
    configs["encounter20k"]["velocities"] = digest(vel)


def main():
    ref = Checker("ref")
    if sys.argv[1:] == ["--only", "neighbours"]:
        write_neighbours(ref)
        return
    if sys.argv[1:] == ["--only", "velocities"]:
        path = os.path.join(HERE, "configs.json")
        with open(path) as f:
            configs = json.load(f)
        add_velocities(ref, configs)
        with open(path, "w") as f:
            json.dump(configs, f, indent=1, sort_keys=True)
        print("wrote velocity digests into configs.json")
        return
    if sys.argv[1:] == ["--only", "encounter2m"]:
        path = os.path.join(HERE, "configs.json")
        with open(path) as f:
            configs = json.load(f)
        configs["encounter2m"] = write_encounter2m(ref)
        with open(path, "w") as f:
            json.dump(configs, f, indent=1, sort_keys=True)
        print("wrote encounter2m.npz and configs.json")
        return
    write_neighbours(ref)
    tiny = {}
    # 1) the reference tests' own small systems
    cases = {
        "mock307": (1, 307, 3, 4, 0.02),     # test_metal_mock.cpp:19-31,72-101 (leaf 4)
        "kernels3000": (0, 3000, 11, 8, 0.01),  # test_gpu_kernels.cpp:29-47
    }
    oracle = Checker("oracle")
    for name, (kind, n, seed, leaf, eps) in cases.items():
        pos, mass = oracle.test_system(kind, n, seed)
        for pack in (0, 1):
            geom, links, order, L, out, info = tree_and_lists(ref, pos, mass, leaf, 0.7, 32,
                                                              pack, eps)
            key = f"{name}_p{pack}"
            tiny[f"{key}_pos"], tiny[f"{key}_mass"] = pos, mass
            tiny[f"{key}_geom"], tiny[f"{key}_links"], tiny[f"{key}_order"] = geom, links, order
            for k in ("groups", "cell_list", "body_list"):
                tiny[f"{key}_{k}"] = L[k]
            tiny[f"{key}_out"] = out
            tiny[f"{key}_eps"] = np.array([eps])
            tiny[f"{key}_leaf"] = np.array([leaf])
    # 2) coincident particles (test_gpu_kernels.cpp:276-288): a leaf bigger than a group
    pos = np.zeros((150, 3))
    pos[:100] = (1.0, 2.0, 3.0)
    i = np.arange(50, dtype=np.float64)
    pos[100:] = np.stack([i, -i, 0.5 * i], 1)
    mass = np.ones(150)
    for pack in (0, 1):
        geom, links, order, L, out, info = tree_and_lists(ref, pos, mass, 8, 0.6, 32, pack, 0.05)
        key = f"coincident_p{pack}"
        tiny[f"{key}_pos"], tiny[f"{key}_mass"] = pos, mass
        tiny[f"{key}_geom"], tiny[f"{key}_links"], tiny[f"{key}_order"] = geom, links, order
        for k in ("groups", "cell_list", "body_list"):
            tiny[f"{key}_{k}"] = L[k]
        tiny[f"{key}_out"] = out
        tiny[f"{key}_eps"] = np.array([0.05])
        tiny[f"{key}_leaf"] = np.array([8])
    np.savez_compressed(os.path.join(HERE, "tiny.npz"), **tiny)

    configs = {}
    # 3) the 50k and 200k galaxies (reference halo profile), theta 0.75, gs 32, packed
    for name, (d, b, h, seed) in {"galaxy50k": (20000, 5000, 25000, 1),
                                  "galaxy200k": (80000, 20000, 100000, 2)}.items():
        pos, vel, mass, typ = ref_galaxy(ref, d, b, h, seed)
        configs[name] = summary(ref, pos, mass)
        configs[name]["model"] = dict(disk_n=d, bulge_n=b, halo_n=h, seed=seed)
        print(name, configs[name]["n_nodes"], configs[name]["n_groups"], flush=True)
    # 4) a small encounter (positions only: the orbit set-up and the second seed)
    pos, vel, mass, typ = ref_encounter(ref, 4000, 1000, 5000, 3)
    configs["encounter20k"] = {"positions": digest(pos), "masses": digest(mass),
                               "model": dict(disk_n=4000, bulge_n=1000, halo_n=5000, seed=3)}
    # 5) leapfrog: the 50k galaxy's positions at rest, 10 steps of 0.01
    pos, vel, mass, typ = ref_galaxy(ref, 20000, 5000, 25000, 1)
    p, v, a, phi = ref.kdk_group(pos, np.zeros_like(pos), mass, 10, 0.01, 1.0, 0.75, 0.025, True,
                                 32, 8, True)
    configs["encounter2m"] = write_encounter2m(ref)
    add_velocities(ref, configs)
    configs["kdk50k"] = {"steps": 10, "dt": 0.01, "pos": digest(p), "vel": digest(v),
                         "acc": digest(a), "pot": digest(phi)}
    with open(os.path.join(HERE, "configs.json"), "w") as f:
        json.dump(configs, f, indent=1, sort_keys=True)
    print("wrote tiny.npz and configs.json")


if __name__ == "__main__":
    main()
