"""Pins the oracle (the SKIPSM restatement in oracle/) against closed forms,
the reference's own scale-constant table, invariants and known answers.

No reference test, golden vector or fixture covers this path (the reference
has no tests for it), so these checks are the pin.
"""
import itertools

import numpy as np
import pytest

import closed_form as cf
from oracle_binding import F32, F64, I16, I32, U8, U16

SIZES_SMALL = list(range(2, 13))
rng0 = np.random.default_rng(1234)
SIZES_RANDOM = [tuple(int(v) for v in rng0.integers(2, 34, size=2)) for _ in range(12)]


def _alpha(rng, h, w, p=0.7):
    return np.where(rng.random((h, w)) < p, 255, 0).astype(np.uint8)


def _data(rng, ptype, h, w, lo=0, hi=32640):
    if ptype == I16:
        return rng.integers(lo, hi + 1, size=(h, w)).astype(np.int16)
    if ptype == I32:
        return rng.integers(lo, 65535 * 128 + 1, size=(h, w)).astype(np.int32)
    # integer-valued doubles: every sum is exact, only the division rounds
    return rng.integers(lo, 1 << 20, size=(h, w)).astype(np.float64)


@pytest.mark.parametrize("ptype", [I16, I32, F64])
@pytest.mark.parametrize("wrap", [False, True])
def test_reduce_alpha_closed_form(oracle, ptype, wrap):
    rng = np.random.default_rng(10 + ptype + 3 * wrap)
    shapes = [(h, w) for h in SIZES_SMALL[:8] for w in SIZES_SMALL[:8]] + SIZES_RANDOM
    for (h, w) in shapes:
        x = _data(rng, ptype, h, w)
        a = _alpha(rng, h, w)
        got, ga = oracle.reduce(x, a, wrap, ptype)
        S, A = cf.reduce_alpha(x, a, wrap)
        if ptype == F64:
            want = np.where(A > 0, S.astype(np.float64) / np.maximum(A, 1).astype(np.float64), 0.0)
        else:
            want = np.where(A > 0, cf.trunc_div(S, np.maximum(A, 1)), 0)
        np.testing.assert_array_equal(got, want.astype(got.dtype), err_msg=f"{h}x{w}")
        np.testing.assert_array_equal(ga, np.where(A > 0, 255, 0).astype(np.uint8))


@pytest.mark.parametrize("ptype", [I16, I32, F64])
@pytest.mark.parametrize("wrap", [False, True])
def test_reduce_plain_closed_form(oracle, ptype, wrap):
    rng = np.random.default_rng(20 + ptype + 3 * wrap)
    shapes = [(h, w) for h in SIZES_SMALL[:8] for w in SIZES_SMALL[:8]] + SIZES_RANDOM
    for (h, w) in shapes:
        x = _data(rng, ptype, h, w)
        got = oracle.reduce(x, None, wrap, ptype)
        S = cf.reduce_plain(x, wrap)
        want = S.astype(np.float64) / 256.0 if ptype == F64 else cf.trunc_div(S, 256)
        np.testing.assert_array_equal(got, want.astype(got.dtype), err_msg=f"{h}x{w}")


def _expand_closed(x, dst, add, wrap, ptype):
    sh, sw = x.shape
    dh, dw = dst.shape
    out = dst.astype(np.float64 if ptype == F64 else np.int64).copy()
    for Y in range(dh):
        for X in range(dw):
            num = cf.expand_sum(x, X, Y, wrap, dw)
            sc = cf.expand_scale(X, Y, sw, sh, wrap, dw)
            e = num / sc if ptype == F64 else int(cf.trunc_div(num, sc))
            out[Y, X] = out[Y, X] + e if add else out[Y, X] - e
    if ptype == I16 and add:
        out = np.clip(out, -32768, 32767)
    return out


@pytest.mark.parametrize("ptype", [I16, I32, F64])
@pytest.mark.parametrize("wrap", [False, True])
@pytest.mark.parametrize("add", [False, True])
def test_expand_closed_form(oracle, ptype, wrap, add):
    rng = np.random.default_rng(30 + ptype + 3 * wrap + 7 * add)
    cases = []
    for sh in range(1, 7):
        for sw in range(1, 7):
            for dh in (2 * sh - 1, 2 * sh):
                for dw in (2 * sw - 1, 2 * sw):
                    if dw < 2 and dh < 2:
                        continue
                    cases.append((sh, sw, dh, dw))
    cases = cases[::3] + [(9, 13, 17, 26), (12, 7, 24, 13), (16, 16, 31, 32)]
    for (sh, sw, dh, dw) in cases:
        x = _data(rng, ptype, sh, sw, lo=-16000, hi=16000) if ptype != F64 else rng.integers(-(1 << 20), 1 << 20, size=(sh, sw)).astype(np.float64)
        d = _data(rng, ptype, dh, dw, lo=-16000, hi=16000) if ptype != F64 else rng.integers(-(1 << 20), 1 << 20, size=(dh, dw)).astype(np.float64)
        got = oracle.expand(x, d, add, wrap, ptype)
        want = _expand_closed(x, d, add, wrap, ptype)
        np.testing.assert_array_equal(got, want.astype(got.dtype), err_msg=f"src {sw}x{sh} dst {dw}x{dh}")


def test_expand_saturates_int16(oracle):
    x = np.full((3, 3), 30000, np.int16)
    d = np.full((6, 6), 30000, np.int16)
    got = oracle.expand(x, d, True, False, I16)
    assert (got == 32767).all()
    got = oracle.expand(x, -d, False, False, I16)   # subtract wraps, no clamp
    assert (got == np.int16(-30000 - 30000 + 65536)).all()


def test_reduce_double_tap_weights(oracle):
    """Non-integer doubles: the oracle's rounding differs from the exact
    closed form by at most a few ulps (pins the tap structure for doubles)."""
    rng = np.random.default_rng(5)
    for wrap in (False, True):
        for (h, w) in [(7, 9), (16, 11), (33, 32)]:
            x = rng.random((h, w)) * 3.0 + 1.0
            a = _alpha(rng, h, w, 0.8)
            got, _ = oracle.reduce(x, a, wrap, F64)
            xs = (x * 2**40).astype(np.int64)  # exact scaled integers
            S, A = cf.reduce_alpha(xs, a, wrap)
            want = np.where(A > 0, (S / 2**40) / np.maximum(A, 1), 0.0)
            np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)


@pytest.mark.parametrize("itype", [U8, U16])
@pytest.mark.parametrize("wrap", [False, True])
def test_collapse_inverts_laplacian(oracle, itype, wrap):
    rng = np.random.default_rng(40 + itype + wrap)
    hi = 255 if itype == U8 else 65535
    for (h, w, L) in [(37, 53, 4), (64, 48, 5), (9, 17, 2), (100, 77, 6)]:
        img = rng.integers(0, hi + 1, size=(h, w, 3)).astype(np.uint8 if itype == U8 else np.uint16)
        a = _alpha(rng, h, w, 0.6)
        lp = oracle.pyramid(img, a, L, wrap, True, itype)
        col = oracle.collapse(lp, wrap, I16 if itype == U8 else I32)
        want = img.transpose(2, 0, 1).astype(np.int64) << 7
        np.testing.assert_array_equal(col[0], want)


def test_constant_is_preserved(oracle):
    for ptype in (I16, I32, F64):
        for wrap in (False, True):
            for (h, w) in [(5, 7), (8, 8), (13, 2), (2, 13)]:
                x = np.full((h, w), 1234, dtype={I16: np.int16, I32: np.int32, F64: np.float64}[ptype])
                got, ga = oracle.reduce(x, np.full((h, w), 255, np.uint8), wrap, ptype)
                assert (got == 1234).all() and (ga == 255).all()
                assert (oracle.reduce(x, None, wrap, ptype) == 1234).all()
                for (dh, dw) in [(2 * h, 2 * w), (2 * h - 1, 2 * w - 1)]:
                    z = np.zeros((dh, dw), x.dtype)
                    assert (oracle.expand(x, z, True, wrap, ptype) == 1234).all(), (ptype, wrap, h, w, dh, dw)


def test_reduce_mask_only_zero(oracle):
    x = np.arange(30, dtype=np.int32).reshape(5, 6)
    got, ga = oracle.reduce(x, np.zeros((5, 6), np.uint8), False, I32)
    assert (got == 0).all() and (ga == 0).all()


def test_reduce_precondition(oracle):
    with pytest.raises(ValueError):
        oracle.reduce(np.zeros((1, 5), np.int32), None, False, I32)
    with pytest.raises(ValueError):
        oracle.reduce(np.zeros((5, 1), np.int32), np.zeros((5, 1), np.uint8), False, I32)


@pytest.mark.parametrize("w,h,levels", [
    (1024, 768, 8), (7168, 3072, 10), (6000, 4000, 10), (2048, 1536, 9),
    (640, 480, 7), (1280, 720, 8), (1920, 1080, 9), (2448, 2048, 9), (3840, 2160, 10), (4096, 3072, 10),
    (8, 100, 1), (9, 9, 2), (17, 17, 3), (1, 1, 1)])
def test_auto_levels(oracle, w, h, levels):
    # src/bounds.h:121-133: halve the short side until it is <= 8.
    s, n = min(w, h), 1
    while n <= 29 and s > 8:
        s = (s + 1) >> 1
        n += 1
    assert n == levels
    assert oracle.auto_levels(w, h) == levels


def test_round_half_away_from_zero(oracle):
    """fromRealPromote known answers through the blend functor: the blend of
    white w and black b at coefficient 1/2 is (w + b)/2 rounded half away
    from zero and clamped."""
    wv = float(255 << 15)
    mask = np.full((1, 4), (255 << 15) // 2, np.int32)  # coefficient just below 1/2
    mask[:] = int(wv / 2)
    white = np.array([[[2, -2, 1, 2147483647]]], np.int32)
    black = np.array([[[3, -3, 2, 2147483647]]], np.int32)
    got = oracle.blend_level(mask, white, black, I32, I32, wv)
    c = float(int(wv / 2)) / wv
    exact = c * white.astype(np.float64) + (1 - c) * black.astype(np.float64)
    want = np.where(exact < 0, np.ceil(exact - 0.5), np.floor(exact + 0.5))
    want = np.clip(want, -2**31, 2**31 - 1)
    np.testing.assert_array_equal(got, want.astype(np.int32))
    # exactly one half: 2.5 -> 3, -2.5 -> -3 (rint would give 2 / -2)
    m = np.full((1, 2), 1 << 14, np.int32)
    got = oracle.blend_level(m, np.array([[[5, -5]]], np.int32), np.array([[[0, 0]]], np.int32), I32, I32, float(1 << 15))
    assert got.tolist() == [[[3, -3]]]
    # Int16 clamps
    m16 = np.full((1, 1), 1 << 6, np.int16)
    got = oracle.blend_level(m16, np.array([[[32767]]], np.int16), np.array([[[32767]]], np.int16), I16, I16, float(1 << 7))
    assert got.tolist() == [[[32767]]]


def test_blend_endpoints(oracle):
    rng = np.random.default_rng(3)
    white = rng.integers(-1000, 1000, size=(3, 4, 5)).astype(np.int32)
    black = rng.integers(-1000, 1000, size=(3, 4, 5)).astype(np.int32)
    wv = float(255 << 15)
    got = oracle.blend_level(np.full((4, 5), 255 << 15, np.int32), white, black, I32, I32, wv)
    np.testing.assert_array_equal(got, white)
    got = oracle.blend_level(np.zeros((4, 5), np.int32), white, black, I32, I32, wv)
    np.testing.assert_array_equal(got, black)


def test_to_from_pyramid_known_answers(oracle):
    img = np.array([[[0, 1, 255]]], np.uint8)
    p = oracle.to_pyramid(img, U8, I16, 7)
    assert p[:, 0, 0].tolist() == [0, 128, 255 * 128]
    # fraction bands: f < 32 -> q; f >= 96 -> q + 1; 32 <= f < 96 -> dither.
    vals = np.array([[[10 * 128 + 31, 10 * 128 + 96, 10 * 128 + 127, -1]]], np.int16).transpose(2, 0, 1)
    mask = np.full((1, 1), 255, np.uint8)
    out = []
    for c in range(4):
        r = oracle.from_pyramid_if(vals[c:c + 1], mask, np.zeros((1, 1, 1), np.uint8), I16, U8, 7, 0, 0)
        out.append(int(r[0, 0, 0]))
    assert out == [10, 11, 11, 0]
    big = np.array([[[32767]]], np.int16)  # q = 255, f = 127 -> 256 -> clamp
    r = oracle.from_pyramid_if(big, mask, np.zeros((1, 1, 1), np.uint8), I16, U8, 7, 0, 0)
    assert r[0, 0, 0] == 255
    # masked-out pixels keep the destination
    r = oracle.from_pyramid_if(big, np.zeros((1, 1), np.uint8), np.full((1, 1, 1), 7, np.uint8), I16, U8, 7, 0, 0)
    assert r[0, 0, 0] == 7


def test_dither_band_statistics(oracle):
    """In the dither band the draw r in [32, 95] rounds up iff r <= f, so the
    round-up rate at fraction f is (f - 31) / 64 (src/fixmath.h:246-264)."""
    n = 20000
    for f in (32, 64, 95):
        v = np.full((1, 1, n), 10 * 128 + f, np.int16)
        r = oracle.from_pyramid_if(v, np.full((1, n), 255, np.uint8), np.zeros((1, n, 1), np.uint8), I16, U8, 7, 1, 99)
        rate = float((r == 11).mean())
        assert abs(rate - (f - 31) / 64) < 0.02, (f, rate)
    draws = {oracle.dither_rand(7, x, 0, 0, 7) for x in range(2000)}
    assert min(draws) == 32 and max(draws) == 95


def test_log_transform_round_trip(oracle):
    rng = np.random.default_rng(8)
    img = np.concatenate([rng.random(50) * 1e2, rng.random(50) * 1e-3, -rng.random(20)]).astype(np.float32).reshape(1, 120, 1)
    p = oracle.to_pyramid(img, F32, F64, 0)
    back = oracle.from_pyramid_if(p, np.full((1, 120), 255, np.uint8), np.zeros((1, 120, 1), np.float32), F64, F32, 0, 0, 0)
    np.testing.assert_allclose(back, img, rtol=2e-7, atol=1e-12)


@pytest.mark.parametrize("itype", [U8, U16])
def test_enblend_pair_mask_extremes(oracle, itype):
    """All-white mask gives the white image, all-black mask the black image,
    wherever the union alpha is set (collapse inverts the Laplacian exactly)."""
    rng = np.random.default_rng(50)
    h, w = 40, 56
    hi = 255 if itype == U8 else 65535
    dt = np.uint8 if itype == U8 else np.uint16
    white = rng.integers(0, hi + 1, size=(h, w, 3)).astype(dt)
    black = rng.integers(0, hi + 1, size=(h, w, 3)).astype(dt)
    wa = _alpha(rng, h, w, 0.7)
    ba = _alpha(rng, h, w, 0.7)
    L = oracle.auto_levels(w, h)
    out, oa = oracle.enblend_pair(white, wa, black, ba, np.full((h, w), 255, np.uint8), L, False, itype, 1, 0)
    union = (wa | ba) != 0
    np.testing.assert_array_equal(oa != 0, union)
    np.testing.assert_array_equal(out[union], white[union])
    out, _ = oracle.enblend_pair(white, wa, black, ba, np.zeros((h, w), np.uint8), L, False, itype, 1, 0)
    np.testing.assert_array_equal(out[union], black[union])
