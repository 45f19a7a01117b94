"""The reference itself, built with the edits of INTEGRATION.md
(oracle/_ref/gt_hip_integration, a binary built in the container from the
reference's sources), running device = hip on the GPU through its own
make_gravity, GroupTreeGravity, DirectGravity and LeapfrogKDK: the grouped
solver on hip builds the same lists as the reference's cpu device (equal
interaction counts) and its forces are within the reference's GPU tolerance of
the cpu device's (test_metal.cpp:115-116: median below 1e-5, max below 1e-3);
the FP32 direct kernel is within test_gpu_kernels.cpp:232-234's bounds of the
FP64 direct sum; ten leapfrog steps on hip stay with the cpu run."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "gt_hip_integration")


@pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref/gt_hip_integration not built "
                    "(make -C oracle integration, where the reference exists)")
def test_reference_runs_device_hip():
    print("running", BIN, "run 50000", flush=True)
    r = subprocess.run([BIN, "run", "50000"], capture_output=True, text=True, timeout=110)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    cases = {d["case"]: d for d in map(json.loads, r.stdout.strip().splitlines())}
    g = cases["group"]
    assert g["summary"] == ("tree with grouped walk (groups of up to 32, packed), single "
                            "precision lists on hip")
    assert g["cells_hip"] == g["cells_cpu"] and g["bodies_hip"] == g["bodies_cpu"]
    assert g["median"] < 1e-5 and g["max"] < 1e-3
    assert g["ms_total"] > 0
    d = cases["direct"]
    assert d["median"] < 1e-5 and d["max"] < 1e-4
    lf = cases["leapfrog"]
    assert lf["max_position_difference"] < 1e-5 * lf["extent"]
