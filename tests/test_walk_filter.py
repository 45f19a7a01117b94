"""The walk's FP32 opening-test filter (galactic-tides_amd/csrc/gt_walk_rec.h)
never decides differently from the exact FP64 test, on random nodes and
groups over 78 octaves of scale and on cases placed within a few ulps of
both thresholds (tests/cpp/test_walk_filter.cpp, built here with g++)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not installed")
def test_filter_agrees_with_exact_test(tmp_path):
    exe = str(tmp_path / "test_walk_filter")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas",
                    "-I", os.path.join(ROOT, "galactic-tides_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "test_walk_filter.cpp"), "-o", exe],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
