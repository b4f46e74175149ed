"""The planning code of the device schedule (lane assignment, descriptors and
workspace layout, start steps, item tables; hsrq3in_amd/csrc/hsrq_sched.cpp)
under AddressSanitizer and UndefinedBehaviorSanitizer on the CPU:
tests/host_plan_checks.cpp, built with g++ (host code only, no GPU), checks
coverage and bounds on random problem sets and context settings."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM_INCLUDE = "/opt/rocm/include"


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir(ROCM_INCLUDE),
                    reason="needs g++ and the ROCm headers")
def test_plan_checks_under_sanitizers(tmp_path):
    exe = str(tmp_path / "host_plan_checks")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-D__HIP_PLATFORM_AMD__", "-I", ROCM_INCLUDE,
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           os.path.join(ROOT, "tests", "host_plan_checks.cpp"),
           os.path.join(ROOT, "hsrq3in_amd", "csrc", "hsrq_sched.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, "200"], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert r.stdout.strip() == "ok 200"
