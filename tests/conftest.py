import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "galactic-tides_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    from pyoracle import Checker
    return Checker("oracle")


@pytest.fixture(scope="session")
def ref():
    from pyoracle import Checker, available
    if not available("ref"):
        pytest.skip("oracle/_ref (the reference built from /root/reference) is not available")
    return Checker("ref")


@pytest.fixture(scope="session")
def hip():
    import gt_hip
    ctx = gt_hip.Context(0)
    print("device:", ctx.device_name, flush=True)
    yield ctx
    ctx.close()


def rel_err(a, b):
    """|a - b| / |b| per row (force error as apps/galactic_tides_bench.cpp:142-144)."""
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
