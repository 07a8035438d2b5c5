"""The C++ layer (include/antimatter_amd/am.hpp) and the BioDynaMo adapter
(include/antimatter_amd/bdm_adapter.hpp), compiled against a mock agent API (tests/adapter/) and
linked to libantimatter_amd.so.  The CPU test only builds the program; the GPU test runs it: the
reference's three free functions through the bridge agree with direct FP64 sums, the store grows
with the agents, a device step annihilates and moves agents, and the standalone BarnesHutTree and
FastMultipoleSolver give the exact sums at theta = 0."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HERE = os.path.join(ROOT, "tests", "adapter")
PKG = os.path.join(ROOT, "antimatter-kernel_amd")


def build(tmp_path):
    exe = str(tmp_path / "adapter_test")
    cmd = ["g++", "-O1", "-std=c++17", "-Wall", "-Wextra", "-I" + os.path.join(ROOT, "include"), "-I" + HERE,
           os.path.join(HERE, "adapter_test.cpp"), "-o", exe, "-L" + PKG, "-lantimatter_amd",
           "-Wl,-rpath," + PKG]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return exe


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_adapter_builds(tmp_path):
    build(tmp_path)


@pytest.mark.gpu
def test_adapter_runs(tmp_path):
    exe = build(tmp_path)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0 and "ADAPTER OK" in res.stdout, res.stdout + res.stderr
