"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

tests/sanitize/host_checks.cpp drives the params.json loader and Validate (host/params.cpp), the
.pvd writer (host/export.cpp) and the CPU oracle's solvers (oracle/cpu_ref.cpp) through ordinary,
malformed, out-of-range and degenerate inputs.  Those three sources are compiled here with g++ and
-fsanitize=address,undefined,float-cast-overflow -fno-sanitize-recover=all, so any heap or stack
overflow, leak, signed overflow, misaligned access or out-of-range float-to-integer cast aborts the
program.  The device library is linked only for the two host helpers the loader calls
(am_species_properties, the error message store); no GPU call is made."""
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT

PKG = os.path.join(ROOT, "antimatter-kernel_amd")
JSON_INC = "/opt/conda/include"


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.skipif(not os.path.exists(os.path.join(JSON_INC, "json.hpp")), reason="needs nlohmann/json")
def test_host_code_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "host_checks")
    srcs = [os.path.join(ROOT, "tests", "sanitize", "host_checks.cpp"), os.path.join(PKG, "host", "params.cpp"),
            os.path.join(PKG, "host", "export.cpp"), os.path.join(ROOT, "oracle", "cpu_ref.cpp")]
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fsanitize=float-cast-overflow", "-fno-sanitize-recover=all",
           "-I" + os.path.join(ROOT, "include"), "-I" + JSON_INC, *srcs, "-o", exe, "-L" + PKG, "-lantimatter_amd",
           "-Wl,-rpath," + PKG]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([exe, os.path.join(GOLDEN, "params_config1.json"), str(tmp_path)], capture_output=True,
                         text=True, timeout=300, env=env)
    out = run.stdout + run.stderr
    assert run.returncode == 0 and "SANITIZE OK" in run.stdout, out
    assert "runtime error" not in out and "ERROR: AddressSanitizer" not in out, out
