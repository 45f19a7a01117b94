"""gt_hip_tree_noalloc, the lists-in evaluation that allocates nothing,
against gt_hip_tree on the same lists (the reference's lists, built by the
oracle): bit for bit in exact and fast mode, on two calls in a row, the
second with a smaller input; and its device-side input check, which must turn
bad indices into GT_HIP_ERR_ARGUMENT before anything is evaluated and leave
the next call unaffected.  The test allocates its device buffers itself
through the HIP runtime (the entry point must not)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class DeviceBuffers:
    """Device memory for the test, through the HIP runtime."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.ptrs = []

    def empty(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(int(nbytes), 16)) == 0
        self.ptrs.append(p.value)
        return p.value

    def put(self, a, slack=0):
        """A device copy of a, with `slack` bytes more allocated after it."""
        a = np.ascontiguousarray(a)
        p = self.empty(a.nbytes + slack)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        self.hip.hipDeviceSynchronize()
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


@pytest.fixture
def dev():
    d = DeviceBuffers()
    yield d
    d.free()


def reference_lists(oracle, kind, n, seed, theta=0.7, gs=32, leaf=8):
    pos, mass = oracle.test_system(kind, n, seed)
    oracle.build(pos, mass, leaf)
    oracle.lists(theta, gs, True)
    return oracle.lists_get()


def call(hip, dev, L, eps, exact, cell_list=None, body_list=None, slack=0):
    import gt_hip
    b, cells = L["bodies"], L["cells"]
    cl = L["cell_list"] if cell_list is None else cell_list
    bl = L["body_list"] if body_list is None else body_list
    g = L["groups"]
    n = len(b)
    p = gt_hip.make_params(n, len(g), 1.0, eps, True)
    d_out = dev.empty(16 * n)
    hip.tree_noalloc(dev.put(b, slack), len(cells), dev.put(cells, slack), len(cl), dev.put(cl),
                     len(bl), dev.put(bl), dev.put(g), p, d_out, exact=exact)
    return dev.get(d_out, (n, 4), np.float32)


@pytest.mark.parametrize("exact", [True, False])
def test_noalloc_equals_tree_on_two_calls(hip, oracle, dev, exact):
    for kind, n, seed, eps in ((1, 20000, 5, 0.01), (0, 3000, 6, 0.02)):  # the second smaller
        L = reference_lists(oracle, kind, n, seed)
        want = hip.tree(L, G=1.0, eps=eps, exact=exact)
        got = call(hip, dev, L, eps, exact)
        print(f"n {n}: {len(L['groups'])} groups, bit-identical "
              f"{np.array_equal(got.view(np.uint32), want.view(np.uint32))}", flush=True)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        if exact:  # the reference's FP32 evaluation (evaluate_group_lists_cpu)
            assert np.array_equal(got.view(np.uint32), oracle.eval(1.0, eps, True).view(np.uint32))


def test_noalloc_rejects_bad_indices_and_recovers(hip, oracle, dev):
    L = reference_lists(oracle, 0, 3000, 7)
    n, nn = len(L["bodies"]), len(L["cells"])
    # one index past the end of each array; the buffers hold one entry more,
    # so nothing outside them could be read even if the check let it through
    cl = L["cell_list"].copy()
    cl[len(cl) // 2] = nn
    with pytest.raises(ValueError, match="cell index out of range"):
        call(hip, dev, L, 0.01, True, cell_list=cl, slack=64)
    bl = L["body_list"].copy()
    bl[len(bl) // 3] = n
    with pytest.raises(ValueError, match="body index out of range"):
        call(hip, dev, L, 0.01, False, body_list=bl, slack=64)
    # the scratch is cleared by every call: a good call right after works
    want = hip.tree(L, G=1.0, eps=0.01, exact=True)
    got = call(hip, dev, L, 0.01, True)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
