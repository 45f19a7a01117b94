"""Benchmark: grouped Barnes-Hut force evaluation on one MI355X.

Workload (one "step" = one leapfrog step, i.e. one full force evaluation -
octree build, group walk, list evaluation, scatter - between FP64 kick/drift
kernels, all on the GPU with the particles resident in HBM):

    parabolic encounter of two Hernquist (1993) galaxies, 1M particles each
    (400k disk, 100k bulge, 500k halo; seeds 3 and 4), N = 2M, second disk
    inclined 60 deg with argument 30 deg, pericentre 2, start separation 30;
    G = 1, Plummer softening 0.025, theta = 0.75, group_size 32 (packed),
    leaf_size 8, quadrupole on, dt = 0.01.

Prints one JSON line: metric value = interactions/s, ms_per_step, the force
error (median and p99 of |a - a_fp64| / |a_fp64| on a seeded sample of 1000
particles against FP64 direct summation), a roofline record for the list-
evaluation kernel and the reference CPU solver timed on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "galactic-tides_amd"))

FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak
HBM_PEAK_GBS = 8000.0     # MI355X published HBM3E bandwidth
CELL_FLOPS, BODY_FLOPS = 71, 21  # per interaction, as written in the reference's kernels
# VALU operations the fast evaluation issues per interaction (its ISA: the
# 32-target batches' loops, one v_rsq_f32 counted as one)
CELL_VALU, BODY_VALU = 33, 16


def encounter_2m(halo="hernquist", ctx=None):
    """The encounter's ICs as galactic-tides-ics makes them (eps 0.025 and
    theta 0.75 measure the realised potential); with a GPU context the
    velocities are measured with its tree (the reference's velocities),
    without one they follow the analytic recipe (positions are the same)."""
    import gt_ics
    m = gt_ics.model(disk_n=400_000, bulge_n=100_000, halo_n=500_000, bulge_mass=0.333,
                     bulge_a=0.2, bulge_rmax=2.0, halo_mass=5.8, halo_rmax=30.0,
                     halo_profile=halo, halo_a=4.0, seed=3, eps=0.025, theta=0.75)
    tree = gt_ics.gpu_tree(ctx) if ctx is not None else None
    return gt_ics.encounter(m, 0.0, 0.0, 60.0, 30.0, 2.0, 30.0, tree=tree)


def galaxy_4m(halo="hernquist", ctx=None):
    import gt_ics
    m = gt_ics.model(disk_n=1_600_000, bulge_n=400_000, halo_n=2_000_000, bulge_mass=0.333,
                     bulge_a=0.2, bulge_rmax=2.0, halo_mass=5.8, halo_rmax=30.0,
                     halo_profile=halo, halo_a=4.0, seed=6, eps=0.025, theta=0.75)
    tree = gt_ics.gpu_tree(ctx) if ctx is not None else None
    pos, vel, mass, _ = gt_ics.galaxy(m, tree=tree)
    return pos, vel, mass


def measured_traffic(kernel="eval_fast"):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a plain copy), or None."""
    path = os.path.join(ROOT, "profiles", "eval_fast_traffic_encounter2m.md")
    try:
        for line in open(path):
            if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                return int(float(line.split(":")[1]))
    except OSError:
        pass
    return None


def host_cpu():
    """The host the CPU baseline ran on: model, logical CPUs, CPUs this process may use."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "cpus_allowed": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "omp_places": os.environ.get("OMP_PLACES")}


def build_hash():
    """SHA-256 (first 16 hex digits) of the HIP library this run loaded."""
    import hashlib
    path = os.path.join(ROOT, "galactic-tides_amd", "libgt_hip.so")
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def cpu_baseline(pos, mass, cfg, offsets=None):
    """The reference's CPU solvers timed on the same particles, on every host
    core with OMP_PROC_BIND=close OMP_PLACES=cores (set here before the OpenMP
    runtime loads, unless the environment sets them): GroupTreeGravity with
    packed lists (the GPU's lists), with unpacked lists (the reference's CPU
    default) and TreeGravity (gravity = tree), each one warm-up then the best
    of 3 compute() calls, as apps/galactic_tides_bench.cpp:128-141 does.  For a
    batch: one packed GroupTreeGravity per system, one warm-up and one call
    each, summed.  The oracle's restatement stands in if the reference build
    is absent."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    host = host_cpu()  # (before the OpenMP runtime binds this thread to one core)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Checker, available
    kind = "reference" if available("ref") else "port"
    chk = Checker("ref" if kind == "reference" else "oracle")
    t0 = time.time()
    args = (cfg["G"], cfg["theta"], cfg["eps"], True, cfg["group_size"], cfg["leaf_size"])
    solvers = {}
    if offsets is None:
        best = chk.time_group(pos, mass, *args, True, repeat=3)
        unpacked = chk.time_group(pos, mass, *args, False, repeat=3)
        solvers = {"group_packed": round(best["total"] * 1e3, 1),
                   "group_unpacked": round(unpacked["total"] * 1e3, 1)}
        if kind == "reference":
            solvers["tree"] = round(chk.time_tree(pos, mass, cfg["G"], cfg["theta"], cfg["eps"],
                                                  True, cfg["leaf_size"], repeat=3) * 1e3, 1)
    else:
        best = None
        for k in range(len(offsets) - 1):
            b, e = int(offsets[k]), int(offsets[k + 1])
            one = chk.time_group(pos[b:e], mass[b:e], *args, True, repeat=1)
            best = one if best is None else {key: best[key] + one[key] for key in best}
        solvers = {"group_packed": round(best["total"] * 1e3, 1)}
    wall = time.time() - t0
    return kind, chk.max_threads(), best, solvers, host, wall


def force_errors(ctx, pos, mass, acc, cfg, offsets=None, size=1000):
    """|a - a64| / |a64| on the reference benchmark's sample of `size` particles
    (std::shuffle of the indices with std::mt19937_64(1),
    apps/galactic_tides_bench.cpp:74-78), a64 by FP64 direct summation over the
    particle's own system (src/gravity.cpp:28-45, on the GPU); median err[n/2]
    and p99 err[n*99/100] as apps/galactic_tides_bench.cpp:142-149."""
    import gt_ics
    n = len(mass)
    sample = np.sort(gt_ics.bench_sample(n, size, 1))
    bounds = np.array([0, n], np.uint64) if offsets is None else np.asarray(offsets, np.uint64)
    err = []
    for k in range(len(bounds) - 1):
        b, e = int(bounds[k]), int(bounds[k + 1])
        mine = sample[(sample >= b) & (sample < e)]
        if len(mine) == 0:
            continue
        ref = ctx.direct_f64_sample(pos[b:e], mass[b:e], mine - b, cfg["G"], cfg["eps"])
        got = acc[mine.astype(np.int64)]
        err.append(np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1))
    err = np.sort(np.concatenate(err))
    return float(err[len(err) // 2]), float(err[len(err) * 99 // 100]), len(err)


def bench_neighbours(args):
    """Octree::build + Octree::for_each_within for every particle of the 2M
    encounter, each with the radius that holds its 32 nearest other particles
    (an SPH smoothing length; found with scipy's k-d tree before timing).  One
    step = one gt_hip_neighbours call; value = queries per second of GPU time
    (build, ordering, walk, scan, moving the lists to their places; HIP
    events), the PCIe-inclusive wall rate alongside."""
    import gt_hip
    from scipy.spatial import cKDTree
    pos, _, mass, _ = encounter_2m(args.halo)
    n = len(mass)
    t0 = time.time()
    radii = cKDTree(pos).query(pos, k=33, workers=-1)[0][:, -1].copy()
    print(f"# radii (33rd nearest, self included) in {time.time() - t0:.1f} s: "
          f"median {np.median(radii):.4f}", file=sys.stderr, flush=True)
    ctx = gt_hip.Context(0)
    total = None
    for _ in range(max(args.warmup, 1)):
        off, idx, d2 = ctx.neighbours(pos, pos, radii, mass=mass, max_out=total)
        total = int(off[-1])
    gpu, wall, phases = [], [], []
    for _ in range(args.steps):
        t = time.perf_counter()
        off, idx, d2 = ctx.neighbours(pos, pos, radii, mass=mass, max_out=total)
        wall.append(time.perf_counter() - t)
        tm = ctx.timings()
        gpu.append(tm["total"])
        phases.append((tm["tree"], tm["lists"], tm["evaluate"]))
    ms = float(np.mean(gpu))
    ph = np.mean(np.array(phases), axis=0)
    print(f"# step (GPU) {ms:.2f} ms: build {ph[0]:.2f}, ordering + walk + scan "
          f"{ph[1]:.2f}, lists to their places {ph[2]:.2f}; wall {1e3 * np.mean(wall):.1f} ms; "
          f"{total} neighbours", file=sys.stderr, flush=True)
    # The walk (each query's first neighbours into a slot of its own, all
    # counted): compulsory bytes = the lists written (8 B index + 8 B squared
    # distance per neighbour), the queries read (24 B centre, 8 B radius) with
    # their order and count (8 B), the tree-order particles (24 B) and the node
    # records (64 B), each once; its phase also holds the ordering and the scan.
    nodes = ctx.info()["n_nodes"]
    traffic = 16 * total + 40 * n + 24 * n + 64 * nodes
    achieved = traffic / (ph[1] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                "frac": round(achieved / 8000.0, 4), "traffic": None, "kernel": "within<2>",
                "kernel_ms": round(float(ph[1]), 4), "bytes_per_launch": traffic,
                "note": "a dependent walk per wave: latency-bound, not bandwidth-bound"}
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from pyoracle import Checker, available
        kind = "reference" if available("ref") else "port"
        chk = Checker("ref" if kind == "reference" else "oracle")
        chk.build(pos, mass, 8)
        m = 200_000
        t = time.perf_counter()
        chk.within(pos[:m], radii[:m])
        el = time.perf_counter() - t
        cpu = {"value": round(m / el, 1), "unit": "queries/s",
               "cores": 1 if kind == "reference" else chk.max_threads(), "kind": kind,
               "sample": f"the first {m} queries of the same workload on the same tree "
                         f"(Octree::for_each_within per query, {el:.1f} s)"}
    line = {
        "metric": "neighbour queries/s (Octree::for_each_within) at N=2M, 32 neighbours each",
        "value": round(n / (ms * 1e-3), 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "wall_ms_per_call": round(1e3 * float(np.mean(wall)), 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64",
        "data": "synthetic: the 2M Hernquist encounter's initial positions",
        "config": {"workload": "neighbours-2M", "n": n, "queries": n, "leaf_size": 8,
                   "radius": "33rd nearest (self included)", "halo": args.halo},
        "neighbours": total,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


# FP64 operations per interaction of Octree::accel_pot as written (sqrt and
# division one each; the opening tests not counted): a cell with quadrupole
# (24 monopole + 49 quadrupole), a particle.
TREE_CELL_FLOPS, TREE_PARTICLE_FLOPS = 73, 24
FP64_PEAK_TFLOPS = 78.6  # MI355X published FP64 vector peak


def bench_tree(args):
    """compute_gravity_tree (Octree::accel_pot per particle, FP64) on the 2M
    encounter's initial positions.  One step = one gt_hip_tree_gravity call
    (tree build + walk); value = interactions per second of GPU time."""
    import gt_hip
    pos, _, mass, _ = encounter_2m(args.halo)
    n = len(mass)
    ctx = gt_hip.Context(0)
    cfg = dict(theta=args.theta, eps=0.025, quadrupole=True, leaf_size=8)
    for _ in range(max(args.warmup, 1)):
        _, _, st = ctx.tree_gravity(pos, mass, **cfg)
    tot, walk, build = [], [], []
    for _ in range(args.steps):
        _, _, st = ctx.tree_gravity(pos, mass, **cfg)
        tm = ctx.timings()
        tot.append(tm["total"])
        walk.append(tm["evaluate"])
        build.append(tm["tree"])
    ms, walk_ms = float(np.mean(tot)), float(np.mean(walk))
    cells, parts = int(st[0]), int(st[1])
    print(f"# step (GPU) {ms:.2f} ms: build {np.mean(build):.2f}, walk {walk_ms:.2f}; "
          f"{cells} cell + {parts} particle interactions", file=sys.stderr, flush=True)
    flops = TREE_CELL_FLOPS * cells + TREE_PARTICLE_FLOPS * parts
    achieved = flops / (walk_ms * 1e-3) / 1e12
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": None,
                "kernel": "tree_gravity", "kernel_ms": round(walk_ms, 4),
                "flops_per_launch": flops, "compute": "fp64 vector ALU"}
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from pyoracle import Checker, available
        if available("ref"):
            chk = Checker("ref")
            best = np.zeros(1)
            t = time.time()
            chk._time_tree(n, np.ascontiguousarray(pos), np.ascontiguousarray(mass), 1.0,
                           args.theta, 0.025, 1, 8, 1, best)
            cpu = {"value": round((cells + parts) / best[0], 1), "unit": "interactions/s",
                   "cores": chk.max_threads(), "kind": "reference",
                   "ms_per_force_eval": round(best[0] * 1e3, 1),
                   "sample": f"the full 2M encounter: TreeGravity::compute (build + walk, "
                             f"OpenMP), one warm-up then one timed call ({time.time() - t:.0f} s)"}
    line = {
        "metric": "FP64 tree-solver interactions/s (gravity = tree) at N=2M",
        "value": round((cells + parts) / (ms * 1e-3), 1),
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64",
        "data": "synthetic: the 2M Hernquist encounter's initial positions",
        "config": {"workload": "tree-2M", "n": n, "theta": args.theta, "eps": 0.025,
                   "leaf_size": 8, "quadrupole": True, "halo": args.halo},
        "interactions": {"cells": cells, "particles": parts},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_direct(args):
    """DirectGravity on the GPU (FP32 O(N^2), gt_direct) for the 200k galaxy
    of the reference's M4 Max figure (README: 122 ms at 200k): one step = one
    gt_hip_direct call; value = pair interactions per second of kernel time."""
    import gt_hip
    import gt_ics
    m = gt_ics.model(disk_n=80_000, bulge_n=20_000, halo_n=100_000, bulge_mass=0.333,
                     bulge_a=0.2, bulge_rmax=2.0, halo_mass=5.8, halo_rmax=30.0,
                     halo_profile=args.halo, halo_a=4.0, seed=2)
    pos, _, mass, _ = gt_ics.galaxy(m)
    n = len(mass)
    centre = 0.5 * (pos.min(0) + pos.max(0))  # bounding_box_centre (gravity_backends.cpp:35-44)
    bodies = np.zeros((n, 4), np.float32)
    bodies[:, :3] = (pos - centre).astype(np.float32)
    bodies[:, 3] = mass.astype(np.float32)
    ctx = gt_hip.Context(0)
    for _ in range(max(args.warmup, 1)):
        out = ctx.direct(bodies, G=1.0, eps=0.025)
    ms = []
    for _ in range(args.steps):
        out = ctx.direct(bodies, G=1.0, eps=0.025)
        ms.append(1e3 * ctx.last_gpu_seconds)
    kms = float(np.mean(ms))
    pairs = n * (n - 1)
    print(f"# direct {n}: {kms:.2f} ms per evaluation (kernel)", file=sys.stderr, flush=True)
    flops = BODY_FLOPS * pairs
    achieved = flops / (kms * 1e-3) / 1e12
    sample = gt_ics.bench_sample(n, 1000, 1)  # apps/galactic_tides_bench.cpp:74-78
    ref = ctx.direct_f64_sample(pos, mass, sample, 1.0, 0.025)
    got = out[sample.astype(np.int64), :3].astype(np.float64)
    err = np.sort(np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1))
    line = {
        "metric": "FP32 direct-summation pair interactions/s (gravity = direct) at N=200k",
        "value": round(pairs / (kms * 1e-3), 1),
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(kms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: a 200k Hernquist galaxy (seed 2)",
        "config": {"workload": "direct-200k", "n": n, "eps": 0.025, "halo": args.halo},
        "force_error": {"median": float(err[len(err) // 2]),
                        "p99": float(err[len(err) * 99 // 100]), "sample": len(err)},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                     "traffic": None, "kernel": "direct_fast", "kernel_ms": round(kms, 4),
                     "flops_per_launch": flops, "compute": "fp32 vector ALU"},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def bench_plain_vs_full(args):
    """The north star's comparison on one GPU, at the 2M encounter and the 4M
    galaxy (theta 0.75, group size 32, eps 0.025): ms per leapfrog step (one
    force evaluation each) of

      plain: the plain GPU group solver, one launch after another with the host
             waiting for each evaluation's status: lists by a count pass and a
             fill pass (predict = 2), groups evaluated in tree order, no split
             lists, no persistent grid, no graphs;
      full:  the learned and batched solver: list slots sized by the trained
             network (50 evaluations gather its samples first, outside the
             timing), every group's work item in one launch on the persistent
             grid taking them longest first, steps from captured HIP graphs;
      uniform: the headline configuration (uniform slots, graphs) for reference.

    All three give bit-identical states (checked here).  Value = plain / full."""
    import gt_hip
    out = {}
    for name, make in (("encounter-2M", lambda c: encounter_2m(args.halo, c)[:3]),
                       ("galaxy-4M", lambda c: galaxy_4m(args.halo, c))):
        ctx = gt_hip.Context(0)
        pos, vel, mass = make(ctx)
        base = dict(G=1.0, theta=0.75, eps=0.025, quadrupole=True, group_size=32, leaf_size=8,
                    pack=1)
        modes = {"plain": (dict(base, predict=2, eval_order=1, split_len=-1, persistent=-1), False),
                 "full": (dict(base, predict=1, persistent=1), True),
                 "uniform": (dict(base, predict=0), True)}
        res, states = {}, {}
        for mode, (cfg, graph) in modes.items():
            if cfg["predict"] == 1:
                ctx.predictor_reset(50)
            ctx.state_set(pos, vel, mass)
            ctx.prime(**cfg)
            ctx.steps(55, 0.01, use_graph=graph, **cfg)  # (trains the network in "full")
            if graph:
                ctx.prepare_steps(args.steps, 0.01, **cfg)
            t = time.perf_counter()
            ctx.steps(args.steps, 0.01, use_graph=graph, **cfg)
            ms = 1e3 * (time.perf_counter() - t) / args.steps
            tm = ctx.timings()
            res[mode] = {"ms_per_step": round(ms, 3), "lists_ms": round(tm["lists"], 3),
                         "evaluate_ms": round(tm["evaluate"], 3), "tree_ms": round(tm["tree"], 3)}
            if mode == "full":
                st = ctx.predictor_stats()
                res[mode]["predictor"] = {k: st[k] for k in ("predicted", "overflows",
                                                             "mean_abs_log_error",
                                                             "slot_entries", "list_entries")}
            states[mode] = ctx.state_get()
            print(f"# {name} {mode}: {res[mode]}", file=sys.stderr, flush=True)
        same = all(np.array_equal(a.view(np.uint64), b.view(np.uint64))
                   for m in ("full", "uniform") for a, b in zip(states[m], states["plain"]))
        out[name] = dict(res, ratio_plain_over_full=round(res["plain"]["ms_per_step"]
                                                          / res["full"]["ms_per_step"], 3),
                         identical_states=bool(same), n=len(mass))
        ctx.close()
    line = {
        "metric": "plain / learned+batched group solver, ms per force evaluation (leapfrog step)",
        "value": out["encounter-2M"]["ratio_plain_over_full"],
        "unit": "x",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 55,
        "ms_per_step": out["encounter-2M"]["full"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (lists, evaluation); fp64 (tree, walk, kick-drift)",
        "data": "synthetic: the 2M Hernquist encounter and the 4M galaxy (seed 6)",
        "config": {"workload": "plain-vs-full", "theta": 0.75, "group_size": 32, "eps": 0.025,
                   "halo": args.halo},
        "results": out,
        "build": {"libgt_hip_sha16": build_hash()},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--halo", default="hernquist", choices=["hernquist", "reference"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--exact", action="store_true", help="bit-exact evaluation mode")
    ap.add_argument("--eval-waves", type=int, default=0, help="tuning: 2 or 3 (0: default)")
    ap.add_argument("--overlap", type=int, default=0,
                    help="tuning: walk ranges evaluated alongside the walk (1 off, 0 default)")
    ap.add_argument("--walk-lanes", type=int, default=0,
                    help="tuning: 1 per-lane walk fetching ahead, 2 plain, 16/32/64 coherent")
    ap.add_argument("--eval-order", type=int, default=0,
                    help="tuning: 1 tree order, 2 longest first per XCD (0: default)")
    ap.add_argument("--split-len", type=int, default=0,
                    help="tuning: list entries per evaluation part (0: default, <0: never split)")
    ap.add_argument("--workload", default="encounter-2m",
                    choices=["encounter-2m", "batch-128", "galaxy-4m", "neighbours-2m",
                             "tree-2m", "direct-200k", "plain-vs-full"],
                    help="encounter-2m: the 2M-particle encounter (headline); batch-128: 128 "
                         "independent galaxies of 2k-100k particles evaluated together; "
                         "galaxy-4m: one 4M-particle galaxy (seed 6); neighbours-2m: "
                         "fixed-radius neighbour queries on the encounter's tree (every "
                         "particle a query, radius = its distance to its 32nd neighbour); "
                         "tree-2m: the FP64 per-particle tree solver (gravity = tree) on the "
                         "encounter; direct-200k: FP32 direct summation (gravity = direct) "
                         "over a 200k-particle galaxy; plain-vs-full: the plain GPU group "
                         "solver against the learned, batched one at 2M and 4M")
    ap.add_argument("--theta", type=float, default=0.75)
    ap.add_argument("--predict", type=int, default=0,
                    help="list slots: 0 uniform, 1 learned predictor (trained on the first "
                         "50 evaluations), 2 two passes (count, fill)")
    ap.add_argument("--eager", action="store_true",
                    help="launch each step's kernels directly instead of from captured graphs")
    ap.add_argument("--hilo", type=int, default=-1,
                    help="hi/lo FP32 positions relative to group centres: 1 on, 0 off "
                         "(default: on for encounter-2m, as its configuration specifies)")
    ap.add_argument("--persistent", type=int, default=-1,
                    help="the persistent evaluation grid (waves taking their XCD's work "
                         "items from a counter): 1 on (default, as the encounter's "
                         "configuration specifies; the autotuner does not change it), 0 off")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default tuning settings (otherwise the autotuner "
                         "picks them on the initial particles before the timed steps)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device 3 times (the plain copy kernel "
                         "that HBM traffic counters are calibrated against), then exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.empty(64 << 20, device="cuda")  # 256 MiB of fp32
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print("copied 3 x 256 MiB", file=sys.stderr)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding "
              "is design-only in this build", file=sys.stderr)
        sys.exit(2)

    if args.workload == "neighbours-2m":
        bench_neighbours(args)
        return
    if args.workload == "tree-2m":
        bench_tree(args)
        return
    if args.workload == "direct-200k":
        bench_direct(args)
        return
    if args.workload == "plain-vs-full":
        bench_plain_vs_full(args)
        return

    import gt_hip
    headline = args.workload == "encounter-2m"
    hilo = bool(args.hilo) if args.hilo >= 0 else headline
    # the library's default is the persistent grid (persistent 0); 0 here asks
    # for one wave per work item (-1)
    persistent = 1 if args.persistent < 0 else (1 if args.persistent else -1)
    cfg = dict(G=1.0, theta=args.theta, eps=0.025, quadrupole=True, group_size=32, leaf_size=8,
               pack=1, exact=args.exact, eval_waves=args.eval_waves, split_len=args.split_len,
               eval_order=args.eval_order, hilo=hilo, predict=args.predict,
               walk_lanes=args.walk_lanes, overlap=args.overlap, persistent=persistent)
    ctx = gt_hip.Context(0)
    print(f"# device: {ctx.device_name}", file=sys.stderr, flush=True)
    t0 = time.time()
    offsets = None
    if args.workload == "batch-128":
        import gt_ics
        offsets, pos, vel, mass = gt_ics.batch_of_galaxies(halo=args.halo)
    elif args.workload == "galaxy-4m":
        pos, vel, mass = galaxy_4m(args.halo, ctx)
    else:
        pos, vel, mass, _ = encounter_2m(args.halo, ctx)
    n = len(mass)
    print(f"# ICs: {n} particles in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    tuning = None
    if not args.no_autotune and args.workload == "encounter-2m" and not args.exact:
        # Only settings that leave the lists alone (group_size stays 32); the
        # persistent grid stays as configured.
        t0 = time.time()
        tuning, tuned_ms = ctx.autotune(pos, mass, **cfg)
        cfg.update({k: v for k, v in tuning.items() if k not in ("group_size", "persistent")})
        tuning["persistent"] = persistent
        print(f"# autotune ({time.time() - t0:.1f} s): {tuning}, {tuned_ms:.2f} ms per evaluation",
              file=sys.stderr, flush=True)
    ctx.state_set(pos, vel, mass)
    if offsets is not None:
        ctx.state_set_systems(offsets)
    ctx.prime(**cfg)
    dt = 0.01
    graph = not args.eager
    if args.warmup:
        ctx.steps(args.warmup, dt, use_graph=graph, **cfg)

    # The K timed steps: one call, which launches them (as one captured graph
    # unless --eager) and synchronises once at the end.  The graph is captured
    # and instantiated beforehand, as set-up: the timed call only launches.
    if graph:
        ctx.prepare_steps(args.steps, dt, **cfg)
    st0 = ctx.pipeline_stats()
    t_start = time.perf_counter()
    ctx.steps(args.steps, dt, use_graph=graph, **cfg)
    elapsed = time.perf_counter() - t_start
    st1 = ctx.pipeline_stats()
    ms_per_step = 1e3 * elapsed / args.steps
    inf = ctx.info()
    tm = ctx.timings()  # per-step means over the timed steps (HIP events)
    interactions = (st1["interactions"] - st0["interactions"]) / args.steps
    value = interactions / (ms_per_step * 1e-3)
    pipeline = {k: st1[k] if k.startswith("last_") else st1[k] - st0[k] for k in st1}
    print(f"# step {ms_per_step:.2f} ms, force eval (GPU events) {tm['total']:.2f} ms: "
          f"tree {tm['tree']:.2f} lists {tm['lists']:.2f} evaluate {tm['evaluate']:.2f} "
          f"scatter+kick {tm['scatter']:.2f}; {inf['n_groups']} groups, {inf['n_nodes']} nodes, "
          f"max list {inf['max_list']}; {pipeline}", file=sys.stderr, flush=True)

    # Force error against FP64 direct summation on a seeded sample of 1000.
    p_now, _, a_now, _ = ctx.state_get()
    med, p99, n_err = force_errors(ctx, p_now, mass, a_now, cfg, offsets)

    # Roofline of the list-evaluation kernel: algorithmic flops of the last
    # evaluation / its kernel time measured with HIP events on the stream it
    # runs on.  It is bound by the FP32 vector ALUs (pairwise gravity has no
    # matrix operation to put on MFMA), priced against the published FP32
    # vector peak; its compulsory HBM bytes (the lists, groups, cells, the
    # targets in and the results out) over the same time give the HBM fraction.
    # With hi/lo positions every cell and body also has a 16 B low part and
    # every group a 32 B hi/lo centre.
    flops = CELL_FLOPS * inf["cell_interactions"] + BODY_FLOPS * inf["body_interactions"]
    eval_ms = tm["evaluate"]
    # Against the GPU's sustained non-packed FP32 VALU rate, measured here
    # (the published 157.3 TFLOP/s needs packed FP32, which this build does
    # not use): the lane operations the kernel issues over its time.
    lane_ops = CELL_VALU * inf["cell_interactions"] + BODY_VALU * inf["body_interactions"]
    valu_peak = ctx.valu_rate()
    valu = {"peak_lane_ops_per_s": round(valu_peak, -9),
            "achieved_lane_ops_per_s": round(lane_ops / (eval_ms * 1e-3), -9),
            "frac": round(lane_ops / (eval_ms * 1e-3) / valu_peak, 4),
            "ops_per_interaction": {"cell": CELL_VALU, "body": BODY_VALU},
            "probe": "16 independent FP32 FMA chains per lane, 8 waves per SIMD"}
    achieved = flops / (eval_ms * 1e-3) / 1e12
    compulsory = (4 * (inf["n_cell_list"] + inf["n_body_list"]) + 32 * inf["n_groups"]
                  + 48 * inf["n_nodes"] + 32 * n)
    if hilo:
        compulsory += 16 * inf["n_nodes"] + 16 * n + 32 * inf["n_groups"]
    hbm_gbs = compulsory / (eval_ms * 1e-3) / 1e9
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                "traffic": measured_traffic() if args.workload == "encounter-2m" else None,
                "kernel": "eval_fast",
                "kernel_ms": round(eval_ms, 4),
                "flops_per_launch": flops, "compute": "fp32 vector ALU",
                "hbm": {"compulsory_bytes": compulsory, "achieved": round(hbm_gbs, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(hbm_gbs / HBM_PEAK_GBS, 4)},
                "valu_measured": valu}

    cpu = None
    if not args.no_cpu:
        kind, cores, best, solvers, host, wall = cpu_baseline(pos, mass, cfg, offsets)
        cpu = {"value": round(best["interactions"] / best["total"], 1), "unit": "interactions/s",
               "cores": cores, "kind": kind,
               "ms_per_force_eval": round(best["total"] * 1e3, 1),
               "split_ms": {k: round(best[k] * 1e3, 1) for k in ("tree", "lists", "evaluate")},
               "solvers_ms": solvers, "host": host,
               "sample": (("the full 2M encounter at the initial positions: GroupTreeGravity on "
                           "the cpu device with packed lists (value; the GPU's lists), with "
                           "unpacked lists, and TreeGravity, each one warm-up then best of 3 "
                           "compute() " if offsets is None else
                           "all 128 galaxies at their initial positions: one GroupTreeGravity "
                           "per galaxy on the cpu device, packed lists, one warm-up then one "
                           "compute() each, summed ") + f"({wall:.0f} s wall)")}

    line = {
        "metric": "ms/force-eval & interactions/s at N=2M theta=0.75; p99 force err vs FP64 direct",
        "value": round(value, 1),
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "ms_per_force_eval": round(tm["total"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (lists, evaluation); fp64 (tree, walk, kick-drift)",
        "data": "synthetic: Hernquist galaxies generated in-process as galactic-tides-ics "
                "generates them (positions; velocities measured with the GPU tree, bit for bit "
                "the reference generator's with its halo; batch-128: analytic velocities)",
        "config": {"workload": {"encounter-2m": "encounter-2M", "batch-128": "batch-128-galaxies",
                                "galaxy-4m": "galaxy-4M"}[args.workload],
                   "systems": 1 if offsets is None else len(offsets) - 1,
                   "n": n, "theta": args.theta, "group_size": 32, "predict": args.predict,
                   "leaf_size": 8, "eps": 0.025, "dt": 0.01, "halo": args.halo,
                   "exact": bool(args.exact), "hilo": hilo, "persistent": persistent > 0,
                   "graphs": graph,
                   "tuning": tuning,
                   "parallelism": "single GPU"},
        "force_error": {"median": med, "p99": p99, "sample": n_err},
        "predictor": ctx.predictor_stats() if args.predict == 1 else None,
        "interactions_per_step": interactions,
        "build": {"libgt_hip_sha16": build_hash()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
