"""Python binding of libgt_hip.so (the C ABI in include/gt_hip.h).

The library is loaded from this directory; if it has not been built this
module raises at import.  There is no CPU fallback: every call runs the HIP
kernels.  Errors from the library become exceptions, following the
reference's conventions: bad arguments raise ValueError (the reference's
std::invalid_argument), device failures raise RuntimeError.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgt_hip.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built: run python galactic-tides_amd/build.py")
_lib = C.CDLL(LIB_PATH)

OK, ERR_ARGUMENT, ERR_DEVICE, ERR_MEMORY, ERR_LIMIT, ERR_UNSUPPORTED = range(6)


class GroupConfig(C.Structure):
    _fields_ = [("G", C.c_double), ("theta", C.c_double), ("eps", C.c_double),
                ("quadrupole", C.c_int), ("group_size", C.c_int), ("leaf_size", C.c_int),
                ("pack", C.c_int), ("exact", C.c_int), ("eval_waves", C.c_int),
                ("walk_lanes", C.c_int), ("split_len", C.c_int), ("eval_order", C.c_int),
                ("hilo", C.c_int), ("predict", C.c_int), ("overlap", C.c_int),
                ("persistent", C.c_int)]


class Params(C.Structure):
    _fields_ = [("n", C.c_uint32), ("n_groups", C.c_uint32), ("G", C.c_float),
                ("eps2", C.c_float), ("quadrupole", C.c_uint32), ("pad0", C.c_uint32),
                ("pad1", C.c_uint32), ("pad2", C.c_uint32)]


class ListsInfo(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("n", "n_nodes", "depth", "n_groups", "n_cell_list",
                                           "n_body_list", "cell_interactions",
                                           "body_interactions", "max_list", "max_group")]


class Timings(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("tree", "lists", "evaluate", "scatter", "total")]


class PredictorStats(C.Structure):
    _fields_ = [("samples", C.c_uint64), ("trained", C.c_uint64), ("train_iterations", C.c_uint64),
                ("train_loss", C.c_double), ("margin", C.c_double), ("predicted", C.c_uint64),
                ("overflows", C.c_uint64), ("mean_abs_log_error", C.c_double),
                ("slot_entries", C.c_uint64), ("list_entries", C.c_uint64)]


class PipelineStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("evaluations", "reruns", "list_rewrites",
                                           "graph_captures", "graph_steps", "graph_stops",
                                           "interactions", "last_exact_tests",
                                           "last_slot_entries", "sort_recounts")]


_vp = C.c_void_p
_P = C.POINTER


def _sig(name, res, args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = args
    return f


_create = _sig("gt_hip_create", C.c_int, [_P(_vp), C.c_int])
_destroy = _sig("gt_hip_destroy", None, [_vp])
_last_error = _sig("gt_hip_last_error", C.c_char_p, [_vp])
_device_name = _sig("gt_hip_device_name", C.c_int, [_vp, C.c_char_p, C.c_size_t])
_last_gpu_seconds = _sig("gt_hip_last_gpu_seconds", C.c_double, [_vp])
_compiled_in = _sig("gt_hip_compiled_in", C.c_int, [])
_direct = _sig("gt_hip_direct", C.c_int, [_vp, _vp, _P(Params), _vp])
_tree = _sig("gt_hip_tree", C.c_int, [_vp, _vp, C.c_uint64, _vp, C.c_uint64, _vp, C.c_uint64,
                                      _vp, _vp, _P(Params), C.c_int, _vp])
_tree_noalloc = _sig("gt_hip_tree_noalloc", C.c_int,
                     [_vp, _vp, C.c_uint64, _vp, C.c_uint64, _vp, C.c_uint64, _vp, _vp,
                      _P(Params), C.c_int, _vp])
_group_compute = _sig("gt_hip_group_compute", C.c_int,
                      [_vp, _P(GroupConfig), C.c_uint64, _vp, _vp, _vp, _vp])
class Diagnostics(C.Structure):
    _fields_ = [("time", C.c_double), ("kinetic", C.c_double), ("potential", C.c_double),
                ("momentum", C.c_double * 3), ("angular_momentum", C.c_double * 3),
                ("centre_of_mass", C.c_double * 3)]


_state_diagnostics = _sig("gt_hip_state_diagnostics", C.c_int,
                          [_vp, C.c_double, C.c_uint64, _vp])
_tree_gravity = _sig("gt_hip_tree_gravity", C.c_int,
                     [_vp, _P(GroupConfig), C.c_uint64, _vp, _vp, _vp, _vp, _vp])
_tree_gravity_at = _sig("gt_hip_tree_gravity_at", C.c_int,
                        [_vp, _P(GroupConfig), C.c_uint64, _vp, _vp, C.c_uint64, _vp, _vp, _vp,
                         _vp])
_neighbours = _sig("gt_hip_neighbours", C.c_int,
                   [_vp, C.c_int, C.c_uint64, _vp, _vp, C.c_uint64, _vp, _vp, C.c_uint64,
                    C.c_uint64, _vp, _vp, _vp, _vp])
_group_compute_batch = _sig("gt_hip_group_compute_batch", C.c_int,
                            [_vp, _P(GroupConfig), C.c_uint64, _vp, _vp, _vp, _vp, _vp])
_state_set_systems = _sig("gt_hip_state_set_systems", C.c_int, [_vp, C.c_uint64, _vp])
_autotune = _sig("gt_hip_autotune", C.c_int, [_vp, _P(GroupConfig), C.c_uint64, _vp, _vp, C.c_int,
                                              _P(GroupConfig), _P(C.c_double)])
_autotune_batch = _sig("gt_hip_autotune_batch", C.c_int,
                       [_vp, _P(GroupConfig), C.c_uint64, _vp, _vp, _vp, C.c_int,
                        _P(GroupConfig), _P(C.c_double)])
_walk_tree = _sig("gt_hip_walk_tree", C.c_int, [_vp, _P(GroupConfig), C.c_uint64, C.c_uint64,
                                                _vp, _vp, _vp, _vp, _vp])
_last_info = _sig("gt_hip_last_info", C.c_int, [_vp, _P(ListsInfo)])
_last_timings = _sig("gt_hip_last_timings", C.c_int, [_vp, _P(Timings)])
_get_tree = _sig("gt_hip_get_tree", C.c_int, [_vp, _vp, _vp, _vp])
_get_lists = _sig("gt_hip_get_lists", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp])
_state_set = _sig("gt_hip_state_set", C.c_int, [_vp, C.c_uint64, _vp, _vp, _vp])
_state_get = _sig("gt_hip_state_get", C.c_int, [_vp, _vp, _vp, _vp, _vp])
_prime = _sig("gt_hip_prime", C.c_int, [_vp, _P(GroupConfig)])
_steps = _sig("gt_hip_steps", C.c_int, [_vp, _P(GroupConfig), C.c_int, C.c_double, C.c_int])
_valu_rate = _sig("gt_hip_debug_valu_rate", C.c_int, [_vp, C.c_int, _P(C.c_double)])
_prepare_steps = _sig("gt_hip_prepare_steps", C.c_int,
                      [_vp, _P(GroupConfig), C.c_int, C.c_double])
_pipeline_stats = _sig("gt_hip_pipeline_stats", C.c_int, [_vp, _P(PipelineStats)])
_predictor_reset = _sig("gt_hip_predictor_reset", C.c_int, [_vp, C.c_int])
_predictor_stats = _sig("gt_hip_predictor_stats", C.c_int, [_vp, _P(PredictorStats)])
_predictor_forward = _sig("gt_hip_predictor_forward", C.c_int,
                          [_vp, C.c_uint64, _vp, _vp, _vp])
_predictor_train_step = _sig("gt_hip_predictor_train_step", C.c_int,
                             [_vp, C.c_uint64, _vp, _vp, _vp, _vp, _vp, C.c_uint32, C.c_uint32,
                              C.c_float, _vp, _vp])
_debug_shrink_slots = _sig("gt_hip_debug_shrink_slots", C.c_int, [_vp, C.c_uint32])
_debug_sort_waits = _sig("gt_hip_debug_sort_waits", C.c_int, [_vp, C.c_uint32])
_direct_f64_sample = _sig("gt_hip_direct_f64_sample", C.c_int,
                          [_vp, C.c_uint64, _vp, _vp, C.c_double, C.c_double, C.c_uint64, _vp,
                           _vp])


def compiled_in():
    return bool(_compiled_in())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_vp)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, np.float64)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def make_config(G=1.0, theta=0.75, eps=0.025, quadrupole=True, group_size=32, leaf_size=8,
                pack=-1, exact=False, eval_waves=0, walk_lanes=0, split_len=0, eval_order=0,
                hilo=False, predict=0, overlap=0, persistent=0):
    cfg = GroupConfig()
    cfg.G, cfg.theta, cfg.eps = G, theta, eps
    cfg.quadrupole, cfg.group_size, cfg.leaf_size = int(quadrupole), group_size, leaf_size
    cfg.pack, cfg.exact, cfg.eval_waves, cfg.walk_lanes = pack, int(exact), eval_waves, walk_lanes
    cfg.split_len, cfg.eval_order, cfg.hilo = split_len, eval_order, int(hilo)
    cfg.predict, cfg.overlap, cfg.persistent = int(predict), int(overlap), int(persistent)
    return cfg


def make_params(n, n_groups=0, G=1.0, eps=0.05, quadrupole=True):
    """make_kernel_params (src/gravity_backends.cpp:25-32)."""
    p = Params()
    p.n, p.n_groups = n, n_groups
    p.G = np.float32(G)
    p.eps2 = np.float32(eps * eps)
    p.quadrupole = int(quadrupole)
    return p


class Context:
    """One GPU context (metal::Context's counterpart).  Not thread-safe."""

    def __init__(self, device=0):
        h = _vp()
        rc = _create(C.byref(h), device)
        if rc != OK:
            msg = _last_error(None).decode()
            raise RuntimeError(f"hip: {msg or 'no GPU device available'} (status {rc})")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc == OK:
            return
        msg = _last_error(self._h).decode()
        if rc == ERR_ARGUMENT:
            raise ValueError(msg)
        if rc == ERR_MEMORY:
            raise MemoryError(msg)
        raise RuntimeError(f"{msg} (status {rc})")

    @property
    def device_name(self):
        buf = C.create_string_buffer(256)
        self._check(_device_name(self._h, buf, 256))
        return buf.value.decode()

    @property
    def last_gpu_seconds(self):
        return _last_gpu_seconds(self._h)

    # -- the reference's GPU entry points ------------------------------------------------
    def direct(self, bodies, G=1.0, eps=0.05):
        bodies = np.ascontiguousarray(bodies, np.float32)
        out = np.empty_like(bodies)
        p = make_params(len(bodies), 0, G, eps)
        self._check(_direct(self._h, _ptr(bodies), C.byref(p), _ptr(out)))
        return out

    def tree(self, lists, G=1.0, eps=0.05, quadrupole=True, exact=False, params=None):
        b = np.ascontiguousarray(lists["bodies"], np.float32)
        cells = np.ascontiguousarray(lists["cells"], np.float32)
        cl = np.ascontiguousarray(lists["cell_list"], np.uint32)
        bl = np.ascontiguousarray(lists["body_list"], np.uint32)
        groups = np.ascontiguousarray(lists["groups"], np.uint32)
        p = params if params is not None else make_params(len(b), len(groups), G, eps, quadrupole)
        out = np.empty((len(b), 4), np.float32)
        self._check(_tree(self._h, _ptr(b), len(cells), _ptr(cells), len(cl), _ptr(cl), len(bl),
                          _ptr(bl), _ptr(groups), C.byref(p), int(exact), _ptr(out)))
        return out

    def tree_noalloc(self, bodies, n_cells, cells, n_cell_list, cell_list, n_body_list,
                     body_list, groups, params, out, exact=False):
        """gt_hip_tree on arrays already in device memory (every array argument
        is a device address, an int); allocates nothing.  params: make_params."""
        self._check(_tree_noalloc(self._h, bodies, n_cells, cells, n_cell_list, cell_list,
                                  n_body_list, body_list, groups, C.byref(params), int(exact),
                                  out))

    # -- the whole force evaluation -----------------------------------------------------
    def group_compute(self, pos, mass, **cfg):
        n = len(mass)
        pos = _f64(pos, (n, 3))
        mass = _f64(mass)
        acc = np.empty((n, 3))
        pot = np.empty(n)
        c = make_config(**cfg)
        self._check(_group_compute(self._h, C.byref(c), n, _ptr(pos), _ptr(mass), _ptr(acc),
                                   _ptr(pot)))
        return acc, pot

    def state_diagnostics(self, time=0.0, systems=1):
        """compute_diagnostics of the resident state: one dict per system."""
        out = (Diagnostics * max(systems, 1))()
        self._check(_state_diagnostics(self._h, time, max(systems, 1), C.cast(out, _vp)))
        return [{"time": d.time, "kinetic": d.kinetic, "potential": d.potential,
                 "momentum": list(d.momentum), "angular_momentum": list(d.angular_momentum),
                 "centre_of_mass": list(d.centre_of_mass)} for d in out]

    def tree_gravity(self, pos, mass, **cfg):
        """The per-particle FP64 tree solver: (acc, pot, [cells, particles])."""
        n = len(mass)
        pos = _f64(pos, (n, 3))
        mass = _f64(mass)
        acc = np.empty((n, 3))
        pot = np.empty(n)
        st = np.zeros(2, np.uint64)
        c = make_config(**cfg)
        self._check(_tree_gravity(self._h, C.byref(c), n, _ptr(pos), _ptr(mass), _ptr(acc),
                                  _ptr(pot), _ptr(st)))
        return acc, pot, st

    def tree_gravity_at(self, pos, mass, points, **cfg):
        """Octree::accel_pot at points (nobody left out) on the tree of the
        particles, FP64: (acc, pot, [cells, particles])."""
        n = len(mass)
        pos = _f64(pos, (n, 3))
        mass = _f64(mass)
        points = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
        npnt = len(points)
        acc = np.empty((npnt, 3))
        pot = np.empty(npnt)
        st = np.zeros(2, np.uint64)
        c = make_config(**cfg)
        self._check(_tree_gravity_at(self._h, C.byref(c), n, _ptr(pos), _ptr(mass), npnt,
                                     _ptr(points), _ptr(acc), _ptr(pot), _ptr(st)))
        return acc, pot, st

    def neighbours(self, pos, centres, radii, mass=None, leaf_size=8, max_out=None):
        """Octree::for_each_within for each query (centre, radius; one radius for
        all or one per query) on the tree of pos: CSR (offsets, index, dist2) in
        the order the reference's callback receives the neighbours.  max_out
        (default: as many as there are, found by a counting call first) bounds
        the lists; when they need more, OverflowError names the total."""
        n = len(pos)
        pos = _f64(pos, (n, 3))
        if mass is not None:
            mass = _f64(mass)
            if len(mass) != n:
                raise ValueError(f"expected {n} masses, got {len(mass)}")
        centres = np.ascontiguousarray(centres, np.float64).reshape(-1, 3)
        nq = len(centres)
        radii = np.ascontiguousarray(np.atleast_1d(radii), np.float64)
        offsets = np.zeros(nq + 1, np.uint64)
        total = C.c_uint64(0)
        mp = _ptr(mass) if mass is not None else None
        if max_out is None:  # count first, then fill
            self._check(_neighbours(self._h, leaf_size, n, _ptr(pos), mp, nq, _ptr(centres),
                                    _ptr(radii), len(radii), 0, _ptr(offsets), None, None,
                                    C.byref(total)))
            max_out = total.value
        index = np.empty(max(max_out, 1), np.uint64)
        dist2 = np.empty(max(max_out, 1))
        rc = _neighbours(self._h, leaf_size, n, _ptr(pos), mp, nq, _ptr(centres), _ptr(radii),
                         len(radii), max_out, _ptr(offsets), _ptr(index), _ptr(dist2),
                         C.byref(total))
        if rc == ERR_LIMIT and total.value > max_out:
            raise OverflowError(f"{total.value} neighbours, room for {max_out}: "
                                + _last_error(self._h).decode())
        self._check(rc)
        t = total.value
        return offsets, index[:t], dist2[:t]

    def group_compute_batch(self, offsets, pos, mass, **cfg):
        """Independent systems: system s is particles offsets[s]:offsets[s+1]."""
        offsets = np.ascontiguousarray(offsets, np.uint64)
        n = int(offsets[-1]) if len(offsets) else 0
        pos = _f64(pos, (n, 3))
        mass = _f64(mass)
        if len(mass) != n:
            raise ValueError(f"expected {n} masses, got {len(mass)}")
        acc = np.empty((n, 3))
        pot = np.empty(n)
        c = make_config(**cfg)
        self._check(_group_compute_batch(self._h, C.byref(c), len(offsets) - 1, _ptr(offsets),
                                         _ptr(pos), _ptr(mass), _ptr(acc), _ptr(pot)))
        return acc, pot

    def autotune(self, pos, mass, search_group_size=False, **cfg):
        """Fastest tuning settings for these particles: (settings dict, ms per evaluation)."""
        n = len(mass)
        pos = _f64(pos, (n, 3))
        mass = _f64(mass)
        c = make_config(**cfg)
        best = GroupConfig()
        ms = C.c_double()
        self._check(_autotune(self._h, C.byref(c), n, _ptr(pos), _ptr(mass),
                              int(search_group_size), C.byref(best), C.byref(ms)))
        keys = ("group_size", "eval_waves", "walk_lanes", "split_len", "eval_order", "persistent")
        return {k: int(getattr(best, k)) for k in keys}, ms.value

    def autotune_batch(self, offsets, pos, mass, search_group_size=False, **cfg):
        """Fastest tuning settings for a batch of independent systems (offsets as
        for group_compute_batch): (settings dict, ms per batched evaluation)."""
        offsets = np.ascontiguousarray(offsets, np.uint64)
        n = int(offsets[-1]) if len(offsets) else 0
        pos = _f64(pos, (n, 3))
        mass = _f64(mass)
        c = make_config(**cfg)
        best = GroupConfig()
        ms = C.c_double()
        self._check(_autotune_batch(self._h, C.byref(c), len(offsets) - 1, _ptr(offsets),
                                    _ptr(pos), _ptr(mass), int(search_group_size),
                                    C.byref(best), C.byref(ms)))
        keys = ("group_size", "eval_waves", "walk_lanes", "split_len", "eval_order", "persistent")
        return {k: int(getattr(best, k)) for k in keys}, ms.value

    def walk_tree(self, geom, links, tree_pos, tree_mass, evaluate=True, **cfg):
        n = len(tree_mass)
        geom = _f64(geom)
        links = np.ascontiguousarray(links, np.int32)
        tpos = _f64(tree_pos, (n, 3))
        tmass = _f64(tree_mass)
        out = np.empty((n, 4), np.float32) if evaluate else None
        c = make_config(**cfg)
        self._check(_walk_tree(self._h, C.byref(c), n, len(links), _ptr(geom), _ptr(links),
                               _ptr(tpos), _ptr(tmass), _ptr(out)))
        return out

    def info(self):
        i = ListsInfo()
        self._check(_last_info(self._h, C.byref(i)))
        return {k: int(getattr(i, k)) for k, _ in ListsInfo._fields_}

    def timings(self):
        t = Timings()
        self._check(_last_timings(self._h, C.byref(t)))
        return {k: float(getattr(t, k)) for k, _ in Timings._fields_}

    def pipeline_stats(self):
        st = PipelineStats()
        self._check(_pipeline_stats(self._h, C.byref(st)))
        return {k: int(getattr(st, k)) for k, _ in PipelineStats._fields_}

    def predictor_reset(self, collect=0):
        self._check(_predictor_reset(self._h, collect))

    def predictor_stats(self):
        st = PredictorStats()
        self._check(_predictor_stats(self._h, C.byref(st)))
        return {k: (float(getattr(st, k)) if t is C.c_double else int(getattr(st, k)))
                for k, t in PredictorStats._fields_}

    PREDICTOR_PARAMS = 4866  # W1 64x8, b1 64, W2 64x64, b2 64, W3 2x64, b3 2

    def predictor_forward(self, x, params):
        """The predictor network's outputs (n x 2, standardised) at normalised
        inputs x (n x 8), computed by the library's inference kernel."""
        x = np.ascontiguousarray(x, np.float32)
        params = np.ascontiguousarray(params, np.float32)
        y = np.empty((len(x), 2), np.float32)
        self._check(_predictor_forward(self._h, len(x), _ptr(x), _ptr(params), _ptr(y)))
        return y

    def predictor_train_step(self, samples, stats, params, m, v, blocks, iteration, lr):
        """One training iteration of the library's kernels: returns (params, m,
        v after the Adam step, per-workgroup gradients (blocks x 4866),
        per-workgroup squared-error sums)."""
        samples = np.ascontiguousarray(samples, np.float32)
        stats = np.ascontiguousarray(stats, np.float32)
        params = np.array(params, np.float32)
        m = np.array(m, np.float32)
        v = np.array(v, np.float32)
        grad = np.empty((blocks, self.PREDICTOR_PARAMS), np.float32)
        loss = np.empty(blocks, np.float32)
        self._check(_predictor_train_step(self._h, len(samples), _ptr(samples), _ptr(stats),
                                          _ptr(params), _ptr(m), _ptr(v), blocks, iteration,
                                          lr, _ptr(grad), _ptr(loss)))
        return params, m, v, grad, loss

    def debug_shrink_slots(self, shrink):
        """Test hook: list slots `shrink` entries too short (0: normal)."""
        self._check(_debug_shrink_slots(self._h, shrink))

    def debug_sort_waits(self, polls):
        """Test hook: radix-sort waits give up after `polls` polls (0: normal)."""
        self._check(_debug_sort_waits(self._h, polls))

    def get_tree(self):
        inf = self.info()
        nn, n = inf["n_nodes"], inf["n"]
        geom = np.empty((nn, 16))
        links = np.empty((nn, 4), np.int32)
        order = np.empty(n, np.uint64)
        self._check(_get_tree(self._h, _ptr(geom), _ptr(links), _ptr(order)))
        return geom, links, order

    def get_lists(self):
        inf = self.info()
        groups = np.empty((inf["n_groups"], 8), np.uint32)
        cl = np.empty(inf["n_cell_list"], np.uint32)
        bl = np.empty(inf["n_body_list"], np.uint32)
        bodies = np.empty((inf["n"], 4), np.float32)
        cells = np.empty((inf["n_nodes"], 12), np.float32)
        origin = np.empty(3)
        self._check(_get_lists(self._h, _ptr(groups), _ptr(cl), _ptr(bl), _ptr(bodies),
                               _ptr(cells), _ptr(origin)))
        return dict(groups=groups, cell_list=cl, body_list=bl, bodies=bodies, cells=cells,
                    origin=origin)

    # -- resident leapfrog ---------------------------------------------------------------
    def state_set(self, pos, vel, mass):
        n = len(mass)
        self._n = n
        self._check(_state_set(self._h, n, _ptr(_f64(pos, (n, 3))), _ptr(_f64(vel, (n, 3))),
                               _ptr(_f64(mass))))

    def state_set_systems(self, offsets):
        offsets = np.ascontiguousarray(offsets, np.uint64)
        self._check(_state_set_systems(self._h, len(offsets) - 1, _ptr(offsets)))

    def state_get(self):
        n = self._n
        pos, vel, acc = (np.empty((n, 3)) for _ in range(3))
        pot = np.empty(n)
        self._check(_state_get(self._h, _ptr(pos), _ptr(vel), _ptr(acc), _ptr(pot)))
        return pos, vel, acc, pot

    def prime(self, **cfg):
        c = make_config(**cfg)
        self._check(_prime(self._h, C.byref(c)))

    def steps(self, nsteps, dt, use_graph=False, **cfg):
        c = make_config(**cfg)
        self._check(_steps(self._h, C.byref(c), nsteps, dt, int(use_graph)))

    def prepare_steps(self, nsteps, dt, **cfg):
        """The graphs steps(nsteps, dt, use_graph=True, **cfg) launches first,
        captured and instantiated now (nothing runs)."""
        c = make_config(**cfg)
        self._check(_prepare_steps(self._h, C.byref(c), nsteps, dt))

    # -- measurement -------------------------------------------------------------------
    def valu_rate(self, rsq=False):
        """Sustained non-packed FP32 VALU rate of the GPU, lane operations per second."""
        r = C.c_double(0.0)
        self._check(_valu_rate(self._h, int(rsq), C.byref(r)))
        return r.value

    def direct_f64_sample(self, pos, mass, sample, G=1.0, eps=0.025):
        n = len(mass)
        sample = np.ascontiguousarray(sample, np.uint64)
        acc = np.empty((len(sample), 3))
        self._check(_direct_f64_sample(self._h, n, _ptr(_f64(pos, (n, 3))), _ptr(_f64(mass)),
                                       G, eps, len(sample), _ptr(sample), _ptr(acc)))
        return acc
