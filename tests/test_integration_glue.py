"""The reference with the edits of INTEGRATION.md applied (device = hip in
include/galactic_tides/gravity_backends.h and src/gravity_backends.cpp):
the edit blocks apply to the reference exactly as written, the patched
reference compiles and links against libgt_hip.so (oracle/_ref/
gt_hip_integration, built by `make -C oracle integration` where the reference
exists), and its make_gravity reports the reference's errors for the new
device (checked here without a GPU; tests/test_gpu_integration.py computes
with it on the GPU)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "gt_hip_integration")
sys.path.insert(0, os.path.join(ROOT, "oracle", "integration"))


def test_every_edit_block_parses():
    import apply_edits
    edits = apply_edits.parse(open(os.path.join(ROOT, "INTEGRATION.md")).read())
    files = {e[0] for e in edits}
    assert files == {"include/galactic_tides/gravity_backends.h", "src/gravity_backends.cpp"}
    assert len(edits) >= 15
    # the three places the reference selects and builds a device
    text = "\n".join(e[3] for e in edits)
    for needle in ("Device::Hip", "dev == \"hip\"", "hip_->group_compute", "hip_->direct"):
        assert needle in text, needle


def test_anchor_rules():
    import apply_edits
    text = "a\nb\na\n"
    assert apply_edits.apply(text, "after-all", "a", "x", "t") == "a\nx\nb\na\nx\n"
    assert apply_edits.apply(text, "replace", "b", "y", "t") == "a\ny\na\n"
    with pytest.raises(ValueError):
        apply_edits.apply(text, "replace", "a", "y", "t")  # twice
    with pytest.raises(ValueError):
        apply_edits.apply(text, "after", "ab", "y", "t")   # not a whole line


@pytest.mark.skipif(not os.path.isdir("/root/reference"), reason="needs the reference sources")
def test_patched_reference_compiles_and_reports_errors():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "integration"], check=True)
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present: the no-GPU errors are checked where there is none")
    out = subprocess.run([BIN, "errors"], capture_output=True, text=True, check=True).stdout
    cases = {d["case"]: d for d in map(json.loads, out.strip().splitlines())}
    assert cases["bad_device"]["result"] == "runtime_error"
    assert cases["bad_device"]["message"] == "device must be 'cpu', 'metal' or 'hip', got 'vulkan'"
    assert cases["group_size_65"] == {"case": "group_size_65", "result": "invalid_argument",
                                      "message": "group_size can be at most 64 on hip"}
    assert cases["tree_on_hip"]["result"] == "runtime_error"
    for c in ("group_on_hip", "direct_on_hip"):
        assert cases[c]["result"] == "runtime_error"
        assert cases[c]["message"] == "hip: no GPU device available"
    assert cases["group_on_cpu"]["result"] == "ok"
