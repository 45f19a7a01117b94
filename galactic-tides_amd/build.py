"""Builds libgt_hip.so for gfx950 in-tree (hipcc, one object per source, then
one shared library).  Run: python galactic-tides_amd/build.py"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgt_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 (v_pk_*_f32) is switched off for every kernel; -O3 would
# otherwise form it from float2-like code.
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-result",
         "-I", os.path.join(ROOT, "include"), "-I", CSRC]
SOURCES = ["gt_api.hip", "gt_build.hip", "gt_walk.hip", "gt_eval.hip", "gt_direct.hip",
           "gt_scan.hip", "gt_sort.hip", "gt_step.hip", "gt_predict.hip", "gt_treegrav.hip", "gt_neighbours.hip",
           "gt_noalloc.hip"]


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    # host library: ICs generator (plain C++, no GPU code)
    host = os.path.join(HERE, "libgt_host.so")
    cmd = ["g++", "-std=c++17", "-O3", "-fPIC", "-shared", "-ffp-contract=off",
           "-I", os.path.join(HERE, "host"), "-o", host + ".tmp",
           os.path.join(HERE, "host", "gt_ics.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host library build failed:\n{r.stderr}")
    os.replace(host + ".tmp", host)
    # C++ test program of the host wrapper (tests/cpp), linked against the library
    exe = os.path.join(HERE, "test_hip_backend")
    cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(HERE, "host"), "-o", exe + ".tmp",
           os.path.join(ROOT, "tests", "cpp", "test_hip_backend.cpp"), "-L", HERE, "-lgt_hip",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ test build failed:\n{r.stderr}")
    os.replace(exe + ".tmp", exe)
    if verbose:
        print("built", LIB, "and", host)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
