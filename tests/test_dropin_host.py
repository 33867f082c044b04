"""Host-side parts of the C++ drop-in that need no GPU: the exposure-weight
plugin interface, check_weight_function and the gray-code table
(include/enblend_amd/exposure_weight.hpp), built and run with g++."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_weight_function_checks(tmp_path):
    exe = tmp_path / "weight_check_test"
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror",
                    os.path.join(ROOT, "tests", "cpp", "weight_check_test.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok  ") == 12
