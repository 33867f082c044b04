"""The batched merge's mask coefficient: the reciprocal product with one
fused residual correction equals the correctly rounded division by 255 for
every value a normalised mask pyramid holds (0 <= v < 2^24), checked
exhaustively in C; the uncorrected product is shown to differ (so the check
can fail)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "coeff_exact.c")


def _run(tmp_path, extra):
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    exe = os.path.join(str(tmp_path), "coeff_exact" + ("_p" if extra else ""))
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", *extra, "-o", exe, SRC, "-lm"], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    return int(out[-1])


def test_fma_quotient_is_correctly_rounded(tmp_path):
    assert _run(tmp_path, []) == 0


def test_plain_product_is_not(tmp_path):
    assert _run(tmp_path, ["-DPRODUCT_ONLY"]) > 0
