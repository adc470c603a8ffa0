"""The C++ host layer compiles on its own (g++, no GPU): SpheralGPU.hh with
only sphgpu.h beside it, and the check program that drives it."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_layer_compiles():
    cmd = ["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "llnl-spheral_amd", "host"),
           os.path.join(ROOT, "tests", "host", "spheral_host_check.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
