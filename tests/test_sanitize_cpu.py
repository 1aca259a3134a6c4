"""CPU: the host C++ of the result records (cistem_downstream_bah_amd/csrc/tm_results.cpp) built with
AddressSanitizer and UndefinedBehaviorSanitizer (g++, no GPU code) and driven by
tests/sanitize/results_harness.cpp: packing records of odd and even sizes with edge cuts and padding,
the master's merge and scaling, the expected threshold, the histogram table and every error path run
without a sanitizer report, and their checks hold."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not available")
def test_result_records_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "results_harness")
    cmd = ["g++", "-std=c++17", "-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "sanitize", "results_harness.cpp"),
           os.path.join(ROOT, "cistem_downstream_bah_amd", "csrc", "tm_results.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120, env=env)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    assert "all checks passed" in r.stdout
