"""CPU: the reference-signature form of the C++ drop-in (include/template_matching_core.hpp) compiles
and links against the library with a caller shaped like match_template's GPU branch
(match_template.cpp:858-985) over stand-in types that carry the reference's member names
(tests/dropin/reference_caller.cpp). No GPU call is made."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_reference_shaped_caller_compiles(tmp_path):
    out = str(tmp_path / "reference_caller")
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"), "-o", out,
                        os.path.join(ROOT, "tests", "dropin", "reference_caller.cpp"),
                        "-L" + os.path.join(ROOT, "cistem_downstream_bah_amd"), "-ltm_hip"],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    assert os.path.getsize(out) > 0
