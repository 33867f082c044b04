"""The float-image transform of the double pyramids
(enblend-enfuse_amd/csrc/logexp.h; src/fixmath.h:165-186, :320-329), built
for the host with g++ and swept against the C library's log1p and expm1:
10.2 million floats over every exponent, their pyramid values back, and a
uniform sweep of [0, 90]."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_log1p_expm1_against_libm(tmp_path):
    exe = tmp_path / "logexp_check"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-o", str(exe),
                    os.path.join(ROOT, "tests", "cpp", "logexp_check.cpp"), "-lm"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True, timeout=120).stdout
    print(out)
    vals = {}
    for line in out.splitlines():
        parts = line.split()
        vals[parts[0]] = {parts[i]: float(parts[i + 1]) for i in range(1, len(parts) - 1, 2)}
    assert vals["log1p"]["max_ulp"] <= 2.0
    assert vals["expm1"]["max_ulp"] <= 3.0
    # the float results of the inverse match the C library's on every sample
    assert vals["expm1"]["float_differences"] == 0
    assert out.strip().endswith("special 1")
