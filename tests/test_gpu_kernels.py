"""GPU parity of the list-evaluation and direct kernels against the CPU oracle
(mirrors tests/test_gpu_kernels.cpp, test_metal_mock.cpp and test_metal.cpp of
the reference)."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_direct_matches_fp64(hip, oracle):
    # test_gpu_kernels.cpp:211-234: N=3000 seed 11, eps 0.01, origin (1,1,1)
    pos, mass = oracle.test_system(0, 3000, 11)
    acc64, pot64 = oracle.direct_f64(pos, mass, 1.0, 0.01)
    bodies = oracle.pack_bodies(pos, mass, np.ones(3))
    out = hip.direct(bodies, 1.0, 0.01)
    e = np.sort(rel_err(out[:, :3].astype(np.float64), acc64))
    ep = np.abs(out[:, 3] / pot64 - 1)
    print("direct vs fp64: median", e[len(e) // 2], "max", e[-1], "pot max", ep.max())
    assert e[len(e) // 2] < 1e-5 and e[-1] < 1e-3 and ep.max() < 1e-5
    cpu = oracle.direct_f32(bodies, 1.0, 0.01)
    d = rel_err(out[:, :3].astype(np.float64), cpu[:, :3].astype(np.float64))
    assert np.median(d) < 1e-5 and d.max() < 1e-3


@pytest.mark.parametrize("n", [1, 63, 5003])
def test_direct_sizes_and_self(hip, oracle, n):
    # a body never feels itself; odd sizes leave the last workgroup partly empty
    pos, mass = oracle.test_system(1, n, 3)
    bodies = oracle.pack_bodies(pos, mass, np.zeros(3))
    out = hip.direct(bodies, 1.0, 0.02)
    if n == 1:
        assert out[0, 0] == 0.0 and out[0, 3] == 0.0
        return
    cpu = oracle.direct_f32(bodies, 1.0, 0.02)
    d = rel_err(out[:, :3].astype(np.float64), cpu[:, :3].astype(np.float64))
    assert d.max() < 1e-4  # summation order and rsqrt differ, nothing else


@pytest.mark.parametrize("pack", [0, 1])
@pytest.mark.parametrize("gs", [8, 32, 64])
def test_tree_exact_is_bitwise(hip, oracle, gs, pack):
    # test_metal_mock.cpp:72-101: the GPU evaluation must equal the CPU's bit for bit
    pos, mass = oracle.test_system(0, 3000, 11)
    oracle.build(pos, mass, 8)
    oracle.lists(0.7, gs, pack)
    L = oracle.lists_get()
    cpu = oracle.eval(1.0, 0.01, True)
    gpu = hip.tree(L, 1.0, 0.01, True, exact=True)
    assert np.array_equal(_bits(gpu), _bits(cpu))


@pytest.mark.parametrize("quad", [True, False])
@pytest.mark.parametrize("gs", [8, 32, 64])
def test_tree_fast_close_to_cpu(hip, oracle, gs, quad):
    # test_metal.cpp:96-119 bounds: median < 1e-5, max < 1e-3 vs the CPU kernel
    pos, mass = oracle.test_system(0, 5003, 5)
    oracle.build(pos, mass, 8)
    oracle.lists(0.5, gs, 1)
    L = oracle.lists_get()
    cpu = oracle.eval(1.0, 0.01, quad)
    gpu = hip.tree(L, 1.0, 0.01, quad, exact=False)
    d = rel_err(gpu[:, :3].astype(np.float64), cpu[:, :3].astype(np.float64))
    dp = np.abs(gpu[:, 3].astype(np.float64) / cpu[:, 3] - 1)
    print("fast vs cpu: median", np.median(d), "max", d.max(), "pot", dp.max())
    assert np.median(d) < 1e-5 and d.max() < 1e-3 and dp.max() < 1e-4


def test_tree_refuses_wide_groups_and_bad_lists(hip, oracle):
    pos, mass = oracle.test_system(0, 500, 2)
    oracle.build(pos, mass, 8)
    oracle.lists(0.7, 32, 1)
    L = oracle.lists_get()
    bad = dict(L)
    g = L["groups"].copy()
    g[0, 1] = g[0, 0] + 65  # 65 targets
    bad["groups"] = g
    with pytest.raises(ValueError):
        hip.tree(bad, 1.0, 0.01)
    bad = dict(L)
    cl = L["cell_list"].copy()
    cl[0] = len(L["cells"])  # out of range
    bad["cell_list"] = cl
    with pytest.raises(ValueError):
        hip.tree(bad, 1.0, 0.01)


def test_cpp_host_wrapper_program():
    # tests/cpp/test_hip_backend.cpp (built by galactic-tides_amd/build.py):
    # the C++ face of the library, checked in the reference's test_metal.cpp style
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "galactic-tides_amd", "test_hip_backend")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
