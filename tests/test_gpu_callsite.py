"""The C++ drop-in (include/enblend_amd/enblend.hpp) driven the way the
reference's callers drive it -- enblendMain's blend step, enfuseMain's mask
generation and fusion loop with a user exposure-weight function shaped like
doc/examples/enfuse/linear.cc, and the single operations through their
argument-object forms (tests/cpp/callsite_test.cpp) -- against the oracle,
bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from oracle_binding import I32, U8, U16, enfuse_params

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "callsite_test")
FWHM = 2.3548200450309493820231386529193992755


@pytest.fixture(scope="module")
def binary():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/_build/callsite_test is not built (run __graft_entry__.build())")
    return BIN


def _run(binary, mode, tmp_path, parts, space=None):
    src, dst = tmp_path / f"{mode}.in", tmp_path / f"{mode}.out"
    with open(src, "wb") as f:
        for p in parts:
            f.write(np.ascontiguousarray(p).tobytes())
    extra = [] if space is None else [str(space)]
    r = subprocess.run([binary, mode, str(src), str(dst)] + extra, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    return open(dst, "rb").read()


def _take(buf, off, dtype, shape):
    n = int(np.prod(shape)) * np.dtype(dtype).itemsize
    return np.frombuffer(buf[off:off + n], dtype).reshape(shape), off + n


@pytest.fixture
def oracle_space(oracle):
    """The oracle's blend colour space, put back to identity afterwards."""
    yield oracle.set_blend_colorspace
    oracle.set_blend_colorspace(0)


@pytest.mark.parametrize("shape,space", [((96, 160), 0), ((301, 457), 0), ((301, 457), 1), ((96, 160), 2),
                                         ((133, 211), 3)])
def test_enblend_call_site_matches_oracle(binary, oracle, oracle_space, tmp_path, shape, space):
    """Also with --blend-colorspace CIELAB / CIELUV / CIECAM (BlendColorspace
    set before the loop)."""
    h, w = shape
    oracle_space(space)
    rng = np.random.default_rng(h + w)
    white = rng.integers(0, 65536, size=(h, w, 3)).astype(np.uint16)
    black = rng.integers(0, 65536, size=(h, w, 3)).astype(np.uint16)
    yy, xx = np.mgrid[0:h, 0:w]
    wa = np.where((xx - 0.3 * w) ** 2 / (0.45 * w) ** 2 + (yy - 0.5 * h) ** 2 / (0.48 * h) ** 2 < 1, 255, 0)
    ba = np.where((xx - 0.7 * w) ** 2 / (0.45 * w) ** 2 + (yy - 0.5 * h) ** 2 / (0.48 * h) ** 2 < 1, 255, 0)
    wa, ba = wa.astype(np.uint8), ba.astype(np.uint8)
    mask = np.where(xx < 0.5 * w + 0.1 * (yy - h / 2), 0, 255).astype(np.uint8)
    L = oracle.auto_levels(w, h)
    print(f"enblend call site {w}x{h} L={L} colour space {space}", flush=True)
    buf = _run(binary, "enblend", tmp_path, [np.array([w, h, L], np.int32), white, wa, black, ba, mask], space)
    got, off = _take(buf, 0, np.uint16, (h, w, 3))
    got_a, off = _take(buf, off, np.uint8, (h, w))
    assert off == len(buf)
    want, want_a = oracle.enblend_pair(white, wa, black, ba, mask, L, False, U16, 1, 0)
    np.testing.assert_array_equal(got_a, want_a)
    np.testing.assert_array_equal(got, want)


def _stack(rng, itype, K, h, w):
    yy, xx = np.mgrid[0:h, 0:w]
    base = 0.5 + 0.4 * np.sin(xx / 7.0) * np.cos(yy / 5.0)
    tint = rng.uniform(0.7, 1.3, size=3)
    m = 255 if itype == U8 else 65535
    out = []
    for k in range(K):
        rad = base[..., None] * tint[None, None, :] * 2.0 ** (k - (K - 1) / 2)
        v = np.clip(rad, 0, 1) ** (1 / 2.2) * m + rng.normal(0, 1, rad.shape)
        out.append(np.clip(np.rint(v), 0, m).astype(np.uint8 if itype == U8 else np.uint16))
    return out


@pytest.mark.parametrize("itype,shape,K,space", [(U8, (64, 96), 3, 0), (U16, (133, 101), 4, 0),
                                                 (U8, (64, 96), 3, 2), (U16, (133, 101), 4, 1),
                                                 (U8, (64, 96), 3, 3), (U16, (133, 101), 4, 3)])
def test_enfuse_call_site_with_user_weight_matches_oracle(binary, oracle, oracle_space, tmp_path, itype, shape, K,
                                                          space):
    h, w = shape
    oracle_space(space)
    rng = np.random.default_rng(31 + itype + K)
    imgs = _stack(rng, itype, K, h, w)
    alphas = [np.where(rng.random((h, w)) < 0.95, 255, 0).astype(np.uint8) for _ in range(K)]
    L = oracle.auto_levels(w, h)
    mode = "enfuse8" if itype == U8 else "enfuse16"
    print(f"enfuse call site {mode} {w}x{h} K={K} L={L} colour space {space}", flush=True)
    buf = _run(binary, mode, tmp_path, [np.array([w, h, L, K], np.int32)] + imgs + alphas, space)
    dt = np.uint8 if itype == U8 else np.uint16
    got, off = _take(buf, 0, dt, (h, w, 3))
    got_a, off = _take(buf, off, np.uint8, (h, w))
    n = 256 if itype == U8 else 65536
    table, off = _take(buf, off, np.float64, (n,))
    assert off == len(buf)
    # the plugin's weights: a triangle of half-width 0.2 * FWHM around 0.5
    z = np.abs((np.arange(n) / (n - 1) - 0.5) / (0.2 * FWHM))
    np.testing.assert_array_equal(table, np.where(z <= 1.0, 1.0 - z, 0.0))
    prm = enfuse_params(w_exposure=1.0, w_saturation=0.2, w_contrast=0.3, contrast_window=5, exposure_table=table)
    want, want_a = oracle.enfuse(imgs, alphas, itype, prm, L, dither=1, seed=0)
    np.testing.assert_array_equal(got_a, want_a)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("shape", [(37, 53), (128, 130)])
def test_single_operations_through_argument_objects(binary, oracle, tmp_path, shape):
    """copyToPyramidImage, reduce with and without alpha, expand subtracting
    and adding, each through its argument-object form."""
    h, w = shape
    rng = np.random.default_rng(7 * h + w)
    img = rng.integers(0, 65536, size=(h, w, 3)).astype(np.uint16)
    a = np.where(rng.random((h, w)) < 0.8, 255, 0).astype(np.uint8)
    print(f"single operations {w}x{h}", flush=True)
    buf = _run(binary, "ops", tmp_path, [np.array([w, h], np.int32), img, a])
    ch, cw = (h + 1) // 2, (w + 1) // 2
    l0, off = _take(buf, 0, np.int32, (h, w, 3))
    r1, off = _take(buf, off, np.int32, (ch, cw, 3))
    a1, off = _take(buf, off, np.uint8, (ch, cw))
    r2, off = _take(buf, off, np.int32, (ch, cw, 3))
    sub, off = _take(buf, off, np.int32, (h, w, 3))
    add, off = _take(buf, off, np.int32, (h, w, 3))
    assert off == len(buf)
    want0 = oracle.to_pyramid(img, U16, I32, 7)
    np.testing.assert_array_equal(l0.transpose(2, 0, 1), want0)
    for c in range(3):
        wr, wa = oracle.reduce(want0[c], a, False, I32)
        np.testing.assert_array_equal(r1[..., c], wr)
        np.testing.assert_array_equal(a1, wa)
        np.testing.assert_array_equal(r2[..., c], oracle.reduce(want0[c], None, False, I32))
        np.testing.assert_array_equal(sub[..., c], oracle.expand(wr, want0[c], False, False, I32))
        np.testing.assert_array_equal(add[..., c], oracle.expand(wr, want0[c], True, False, I32))
