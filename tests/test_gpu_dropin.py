"""The C++ drop-in templates (include/enblend_amd/enblend.hpp), instantiated
with the reference's template arguments for 16-bit RGB images, produce the
oracle's Laplacian pyramid and collapse it back exactly."""
import os
import subprocess

import numpy as np
import pytest

from oracle_binding import U16

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "dropin_test")


@pytest.mark.parametrize("w,h", [(97, 61), (640, 480)])
def test_dropin_laplacian_and_collapse(oracle, tmp_path, w, h):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/_build/dropin_test is not built (run __graft_entry__.build())")
    rng = np.random.default_rng(w + h)
    img = rng.integers(0, 65536, size=(h, w, 3)).astype(np.uint16)
    a = np.where(rng.random((h, w)) < 0.85, 255, 0).astype(np.uint8)
    L = oracle.auto_levels(w, h)
    src = tmp_path / "in.bin"
    dst = tmp_path / "out.bin"
    with open(src, "wb") as f:
        f.write(np.array([w, h, L], np.int32).tobytes())
        f.write(img.tobytes())
        f.write(a.tobytes())
    print(f"dropin {w}x{h} L={L}", flush=True)
    r = subprocess.run([BIN, str(src), str(dst)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(dst, np.int32)
    want = oracle.pyramid(img, a, L, False, True, U16)
    off = 0
    for l, lv in enumerate(want):
        n = lv.size
        np.testing.assert_array_equal(got[off:off + n].reshape(lv.shape), lv, err_msg=f"level {l}")
        off += n
    assert off == got.size
