"""ctypes access to the CPU checkers (oracle/gt_oracle.h).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg; the product library never uses it.

    Checker("oracle")  -> oracle/libgt_oracle.so   (restatement of the reference)
    Checker("ref")     -> oracle/_ref/libgt_ref.so  (the reference's own sources)
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATHS = {"oracle": os.path.join(HERE, "libgt_oracle.so"),
         "ref": os.path.join(HERE, "_ref", "libgt_ref.so")}
PREFIX = {"oracle": "gto_", "ref": "gtref_"}

_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u32 = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64 = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_u64t = C.c_uint64
_d = C.c_double
_i = C.c_int

_SIGS = {
    "new": (_vp, []),
    "free": (None, [_vp]),
    "build": (_i, [_vp, _u64t, _dp, _dp, _i, _u64t]),
    "lists": (_i, [_vp, _d, _i, _i]),
    "info": (None, [_vp, _u64]),
    "nodes": (None, [_vp, _dp, _i32]),
    "order": (None, [_vp, _u64]),
    "tree_positions": (None, [_vp, _dp, _dp]),
    "lists_get": (None, [_vp, _u32, _u32, _u32, _fp, _fp, _dp]),
    "eval": (None, [_vp, _d, _d, _i, _fp]),
    "tree_gravity": (None, [_vp, _u64t, _dp, _dp, _d, _d, _d, _i, _i, _dp, _dp, _u64]),
    "accel_at": (None, [_vp, _u64t, _dp, _d, _d, _d, _i, _dp, _dp]),
    "direct_f64": (None, [_u64t, _dp, _dp, _d, _d, _dp, _dp]),
    "direct_f64_sample": (None, [_u64t, _dp, _dp, _d, _d, _u64t, _u64, _dp]),
    "direct_f32": (None, [_u64t, _fp, _d, _d, _fp]),
    "pack_bodies": (None, [_u64t, _dp, _dp, _dp, _fp]),
    "group_gravity": (_i, [_vp, _u64t, _dp, _dp, _d, _d, _d, _i, _i, _i, _i, _dp, _dp]),
    "kdk_group": (_i, [_vp, _u64t, _dp, _dp, _dp, _d, _d, _d, _i, _i, _i, _i, _i, _d, _dp, _dp]),
    "time_group": (_i, [_vp, _u64t, _dp, _dp, _d, _d, _d, _i, _i, _i, _i, _i, _dp]),
    "max_threads": (_i, []),
    "within": (_u64t, [_vp, _u64t, _dp, _dp, _u64t, _u64, _u64t, _vp, _vp]),
}
_ORACLE_ONLY = {"test_system": (None, [_i, _u64t, _u64t, _dp, _dp])}
_REF_ONLY = {
    "galaxy": (_u64t, [_dp, _d, _d, _dp, _dp, _dp, _u8]),
    "encounter": (_u64t, [_dp, _d, _d, _d, _d, _d, _d, _dp, _dp, _dp, _u8]),
    "time_tree": (_i, [_u64t, _dp, _dp, _d, _d, _d, _i, _i, _i, _dp]),
    "bench_sample": (_u64t, [_u64t, _u64t, _u64t, _u64]),
}

INFO_KEYS = ("n", "n_nodes", "depth", "n_groups", "n_cell_list", "n_body_list",
             "cell_interactions", "body_interactions", "max_list", "max_group")


def available(kind):
    return os.path.exists(PATHS[kind])


class Checker:
    """One CPU checker library; a `run` handle holds a tree and its lists."""

    def __init__(self, kind="oracle"):
        if not available(kind):
            raise FileNotFoundError(f"{PATHS[kind]} is not built (run oracle/Makefile)")
        self.kind = kind
        self.lib = C.CDLL(PATHS[kind])
        sigs = dict(_SIGS)
        sigs.update(_ORACLE_ONLY if kind == "oracle" else _REF_ONLY)
        for name, (res, args) in sigs.items():
            f = getattr(self.lib, PREFIX[kind] + name)
            f.restype = res
            f.argtypes = args
            setattr(self, "_" + name, f)
        self.h = self._new()

    def __del__(self):
        try:
            self._free(self.h)
        except Exception:
            pass

    # -- tree and lists ---------------------------------------------------------
    def build(self, pos, mass, leaf_size=8, parallel_threshold=1 << 40):
        pos = np.ascontiguousarray(pos, np.float64)
        mass = np.ascontiguousarray(mass, np.float64)
        rc = self._build(self.h, len(mass), pos, mass, leaf_size, parallel_threshold)
        if rc:
            raise RuntimeError("build failed")

    def lists(self, theta, group_size=32, pack=True):
        if self._lists(self.h, theta, group_size, int(pack)):
            raise ValueError("group_size must be at least 1")

    def info(self):
        a = np.zeros(10, np.uint64)
        self._info(self.h, a)
        return {k: int(v) for k, v in zip(INFO_KEYS, a)}

    def nodes(self):
        nn = self.info()["n_nodes"]
        geom = np.zeros((nn, 16), np.float64)
        links = np.zeros((nn, 4), np.int32)
        self._nodes(self.h, geom, links)
        return geom, links

    def order(self):
        o = np.zeros(self.info()["n"], np.uint64)
        self._order(self.h, o)
        return o

    def tree_positions(self):
        n = self.info()["n"]
        p = np.zeros((n, 3), np.float64)
        m = np.zeros(n, np.float64)
        self._tree_positions(self.h, p, m)
        return p, m

    def lists_get(self):
        inf = self.info()
        groups = np.zeros((inf["n_groups"], 8), np.uint32)
        cl = np.zeros(max(inf["n_cell_list"], 1), np.uint32)
        bl = np.zeros(max(inf["n_body_list"], 1), np.uint32)
        bodies = np.zeros((inf["n"], 4), np.float32)
        cells = np.zeros((inf["n_nodes"], 12), np.float32)
        origin = np.zeros(3, np.float64)
        self._lists_get(self.h, groups, cl, bl, bodies, cells, origin)
        return dict(groups=groups, cell_list=cl[:inf["n_cell_list"]],
                    body_list=bl[:inf["n_body_list"]], bodies=bodies, cells=cells,
                    origin=origin)

    def eval(self, G=1.0, eps=0.05, quadrupole=True):
        out = np.zeros((self.info()["n"], 4), np.float32)
        self._eval(self.h, G, eps, int(quadrupole), out)
        return out

    # -- whole solvers ------------------------------------------------------------
    def tree_gravity(self, pos, mass, G=1.0, theta=0.6, eps=0.05, quadrupole=True, leaf_size=8):
        n = len(mass)
        acc = np.zeros((n, 3))
        pot = np.zeros(n)
        st = np.zeros(2, np.uint64)
        self._tree_gravity(self.h, n, np.ascontiguousarray(pos, np.float64),
                           np.ascontiguousarray(mass, np.float64), G, theta, eps,
                           int(quadrupole), leaf_size, acc, pot, st)
        return acc, pot, st

    def accel_at(self, points, G=1.0, theta=0.6, eps=0.05, quadrupole=True):
        """Octree::accel_pot at points (nobody left out) on the last built tree."""
        points = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
        acc = np.zeros((len(points), 3))
        pot = np.zeros(len(points))
        self._accel_at(self.h, len(points), points, G, theta, eps, int(quadrupole), acc, pot)
        return acc, pot

    def direct_f64(self, pos, mass, G=1.0, eps=0.05):
        n = len(mass)
        acc = np.zeros((n, 3))
        pot = np.zeros(n)
        self._direct_f64(n, np.ascontiguousarray(pos, np.float64),
                         np.ascontiguousarray(mass, np.float64), G, eps, acc, pot)
        return acc, pot

    def direct_f64_sample(self, pos, mass, sample, G=1.0, eps=0.05):
        sample = np.ascontiguousarray(sample, np.uint64)
        acc = np.zeros((len(sample), 3))
        self._direct_f64_sample(len(mass), np.ascontiguousarray(pos, np.float64),
                                np.ascontiguousarray(mass, np.float64), G, eps, len(sample),
                                sample, acc)
        return acc

    def direct_f32(self, bodies, G=1.0, eps=0.05):
        bodies = np.ascontiguousarray(bodies, np.float32)
        out = np.zeros_like(bodies)
        self._direct_f32(len(bodies), bodies, G, eps, out)
        return out

    def pack_bodies(self, pos, mass, origin):
        out = np.zeros((len(mass), 4), np.float32)
        self._pack_bodies(len(mass), np.ascontiguousarray(pos, np.float64),
                          np.ascontiguousarray(mass, np.float64),
                          np.ascontiguousarray(origin, np.float64), out)
        return out

    def group_gravity(self, pos, mass, G=1.0, theta=0.75, eps=0.025, quadrupole=True,
                      group_size=32, leaf_size=8, pack=True):
        n = len(mass)
        acc = np.zeros((n, 3))
        pot = np.zeros(n)
        rc = self._group_gravity(self.h, n, np.ascontiguousarray(pos, np.float64),
                                 np.ascontiguousarray(mass, np.float64), G, theta, eps,
                                 int(quadrupole), group_size, leaf_size, int(pack), acc, pot)
        if rc:
            raise RuntimeError("group gravity failed")
        return acc, pot

    def kdk_group(self, pos, vel, mass, nsteps, dt, G=1.0, theta=0.75, eps=0.025,
                  quadrupole=True, group_size=32, leaf_size=8, pack=True):
        n = len(mass)
        pos = np.array(pos, np.float64, order="C")
        vel = np.array(vel, np.float64, order="C")
        acc = np.zeros((n, 3))
        pot = np.zeros(n)
        rc = self._kdk_group(self.h, n, pos, vel, np.ascontiguousarray(mass, np.float64), G,
                             theta, eps, int(quadrupole), group_size, leaf_size, int(pack),
                             nsteps, dt, acc, pot)
        if rc:
            raise RuntimeError("kdk failed")
        return pos, vel, acc, pot

    def time_group(self, pos, mass, G=1.0, theta=0.75, eps=0.025, quadrupole=True,
                   group_size=32, leaf_size=8, pack=True, repeat=3):
        best = np.zeros(6)
        rc = self._time_group(self.h, len(mass), np.ascontiguousarray(pos, np.float64),
                              np.ascontiguousarray(mass, np.float64), G, theta, eps,
                              int(quadrupole), group_size, leaf_size, int(pack), repeat, best)
        if rc:
            raise RuntimeError("timing failed")
        return dict(total=best[0], tree=best[1], lists=best[2], evaluate=best[3],
                    interactions=best[4] + best[5])

    def time_tree(self, pos, mass, G=1.0, theta=0.75, eps=0.025, quadrupole=True, leaf_size=8,
                  repeat=3):
        """(reference build only) TreeGravity::compute wall time: one warm-up,
        then the best of `repeat` calls, in seconds."""
        best = np.zeros(1)
        rc = self._time_tree(len(mass), np.ascontiguousarray(pos, np.float64),
                             np.ascontiguousarray(mass, np.float64), G, theta, eps,
                             int(quadrupole), leaf_size, repeat, best)
        if rc:
            raise RuntimeError("timing failed")
        return float(best[0])

    def within(self, centres, radii):
        """Octree::for_each_within for each query on the last build's tree: CSR
        offsets (nq + 1), original indices and squared distances, in the order
        the reference's callback receives them."""
        centres = np.ascontiguousarray(centres, np.float64).reshape(-1, 3)
        radii = np.ascontiguousarray(np.atleast_1d(radii), np.float64)
        nq = len(centres)
        offsets = np.zeros(nq + 1, np.uint64)
        total = self._within(self.h, nq, centres, radii, len(radii), offsets, 0, None, None)
        index = np.zeros(max(total, 1), np.uint64)
        dist2 = np.zeros(max(total, 1))
        self._within(self.h, nq, centres, radii, len(radii), offsets, total,
                     index.ctypes.data, dist2.ctypes.data)
        return offsets, index[:total], dist2[:total]

    def max_threads(self):
        return int(self._max_threads())

    def bench_sample(self, n, size=1000, seed=1):
        """(reference build only) The benchmark's force-error sample,
        apps/galactic_tides_bench.cpp:74-78."""
        out = np.zeros(min(n, size), np.uint64)
        self._bench_sample(n, seed, size, out)
        return out

    # -- inputs -------------------------------------------------------------------
    def test_system(self, kind, n, seed):
        pos = np.zeros((n, 3))
        mass = np.zeros(n)
        self._test_system(kind, n, seed, pos, mass)
        return pos, mass


def canonical_ids(links):
    """Preorder index of every node (children in stored order): a numbering
    that does not depend on how a builder laid its nodes out."""
    nn = len(links)
    canon = np.full(nn, -1, np.int64)
    if nn == 0:
        return canon
    stack = [0]
    k = 0
    while stack:
        ni = stack.pop()
        canon[ni] = k
        k += 1
        fc, nc = int(links[ni, 0]), int(links[ni, 1])
        for c in range(nc - 1, -1, -1):
            stack.append(fc + c)
    return canon
