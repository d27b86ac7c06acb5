"""The grouped-batch planner (codyfortran_amd/csrc/batch_plan.cpp), host-only C++, built with
AddressSanitizer and UBSan and run on the CPU: known block counts of the 4096-point system, full
coverage of every batched shape, refused shapes, tampered maps and the 2^31 overflow guard."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "codyfortran_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_batch_plan_under_sanitizers(tmp_path):
    exe = str(tmp_path / "batch_plan_test")
    cmd = ["g++", "-std=c++17", "-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-I", CSRC, os.path.join(ROOT, "tests", "cpu", "batch_plan_test.cpp"),
           os.path.join(CSRC, "batch_plan.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all passed" in r.stdout
