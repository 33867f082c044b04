"""Pins the oracle's enfuse weights and fusion (src/enfuse.h) by direct
per-pixel recomputation from the formulas, exact-integer arithmetic where
the reference's arithmetic is exact, and known answers."""
import math
from collections import Counter

import numpy as np
import pytest

from oracle_binding import (EXP_BISQUARE, EXP_FULLSINE, EXP_GAUSSIAN, EXP_HALFSINE, EXP_LORENTZIAN, F32, U8, U16,
                            enfuse_params)

FWHM = {EXP_GAUSSIAN: 2.3548200450309493820231386529193992755, EXP_LORENTZIAN: 2.8284271247461900976033774484193961571,
        EXP_HALFSINE: 2.0943951023931954923084289221863352561, EXP_FULLSINE: math.pi,
        EXP_BISQUARE: 1.6817928305074290860622509524664297901}
MAXV = {U8: 255.0, U16: 65535.0}


def round_half_away(v):
    return math.floor(v + 0.5) if v >= 0 else math.ceil(v - 0.5)


def gray_int(px, itype):
    v = (float(px[0]) + float(px[1]) + float(px[2])) / 3.0
    return int(min(MAXV[itype], max(0.0, round_half_away(v))))


def weight_fn(fn, opt, width, y):
    width = width * FWHM[EXP_GAUSSIAN] / FWHM[fn] if fn else width
    z = (y - opt) / width
    if fn == EXP_GAUSSIAN:
        return math.exp(-0.5 * z * z)
    if fn == EXP_LORENTZIAN:
        return 1.0 / (1.0 + 0.5 * z * z)
    if fn == EXP_HALFSINE:
        return math.cos(z) if abs(z) <= math.pi / 2 else 0.0
    if fn == EXP_FULLSINE:
        return 0.5 + math.cos(z) / 2.0 if abs(z) <= math.pi else 0.0
    if abs(z) <= 1.0:
        return (1.0 - z * z) * (1.0 + z * z)
    return 0.0


def f32(v):
    return float(np.float32(v))


def direct_mask(img, alpha, itype, w_exp=0.0, w_sat=0.0, w_con=0.0, w_ent=0.0, fn=EXP_GAUSSIAN, opt=0.5, width=0.2,
                cwin=5, ewin=3):
    """Per-pixel restatement of enfuseMask for integer RGB images."""
    h, w, _ = img.shape
    maxv = MAXV[itype]
    out = np.zeros((h, w), np.float32)
    pl = [math.log(i) if i else 0.0 for i in range(ewin * ewin + 1)]
    pe = [(i / (ewin * ewin)) * math.log(i / (ewin * ewin)) if i else 0.0 for i in range(ewin * ewin + 1)]
    for y in range(h):
        for x in range(w):
            if not alpha[y, x]:
                continue
            m = np.float32(0.0)
            px = img[y, x]
            if w_exp > 0:
                g = gray_int(px, itype)
                m = np.float32(w_exp * weight_fn(fn, opt, width, g / maxv))
            if w_con > 0:
                b = cwin // 2
                sd = 0
                if b <= y < h - b and b <= x < w - b:
                    vals = [gray_int(img[yy, xx], itype) for yy in range(y - b, y + b + 1)
                            for xx in range(x - b, x + b + 1) if alpha[yy, xx]]
                    n = len(vals)
                    if n > 1:
                        s, q = float(sum(vals)), float(sum(v * v for v in vals))
                        sd = round_half_away(math.sqrt((q - s * s / n) / (n - 1)))
                m = np.float32(f32(w_con * sd / maxv) + m)
            if w_sat > 0:
                mx, mn = max(int(c) for c in px), min(int(c) for c in px)
                sat = 0.0
                if mx != mn:
                    sm, df = float(mx + mn), float(mx - mn)
                    sat = f32(w_sat * (df / sm if sm <= maxv else df / (2.0 * maxv - sm)))
                m = np.float32(np.float32(sat) + m)
            if w_ent > 0:
                b = ewin // 2
                ents = [0, 0, 0]
                if b <= y < h - b and b <= x < w - b:
                    for c in range(3):
                        vals = [int(img[yy, xx, c]) for yy in range(y - b, y + b + 1)
                                for xx in range(x - b, x + b + 1) if alpha[yy, xx]]
                        cnt = Counter(vals)
                        total, bins = len(vals), len(cnt)
                        e = 0.0
                        if total and bins > 1:
                            if total == ewin * ewin:
                                for k in sorted(cnt):
                                    e += pe[cnt[k]]
                                e = -e / pl[bins]
                            else:
                                for k in sorted(cnt):
                                    p = cnt[k] / float(total)
                                    e += p * math.log(p)
                                e = -e / math.log(bins)
                        ents[c] = int(min(maxv, max(0, round_half_away(e * maxv))))
                m = np.float32(f32(w_ent * float(np.float32(min(ents)))) + m)
            out[y, x] = m
    return out


def _img(rng, itype, h, w, levels=None):
    hi = 255 if itype == U8 else 65535
    dt = np.uint8 if itype == U8 else np.uint16
    if levels:   # few distinct values so histograms have repeated keys
        return rng.choice(np.linspace(0, hi, levels).astype(dt), size=(h, w, 3))
    return rng.integers(0, hi + 1, size=(h, w, 3)).astype(dt)


@pytest.mark.parametrize("itype", [U8, U16])
@pytest.mark.parametrize("fn", [EXP_GAUSSIAN, EXP_LORENTZIAN, EXP_HALFSINE, EXP_FULLSINE, EXP_BISQUARE])
def test_exposure_weight(oracle, itype, fn):
    rng = np.random.default_rng(fn + itype)
    img = _img(rng, itype, 11, 13)
    a = np.where(rng.random((11, 13)) < 0.8, 255, 0).astype(np.uint8)
    prm = enfuse_params(w_exposure=0.7, exposure_function=fn, exposure_optimum=0.4, exposure_width=0.25)
    got = oracle.enfuse_mask(img, a, itype, prm)
    want = direct_mask(img, a, itype, w_exp=0.7, fn=fn, opt=0.4, width=0.25)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("itype", [U8, U16])
def test_exposure_table_matches_builtin(oracle, itype):
    """A tabulated ExposureWeight (the user-plugin path) gives the built-in's
    weights bit for bit when the table holds weight(code / max)."""
    rng = np.random.default_rng(3)
    img = _img(rng, itype, 9, 10)
    a = np.full((9, 10), 255, np.uint8)
    maxv = MAXV[itype]
    tab = np.array([weight_fn(EXP_GAUSSIAN, 0.5, 0.2, g / maxv) for g in range(int(maxv) + 1)])
    got = oracle.enfuse_mask(img, a, itype, enfuse_params(exposure_table=tab))
    want = oracle.enfuse_mask(img, a, itype, enfuse_params())
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("itype", [U8, U16])
def test_all_weights(oracle, itype):
    rng = np.random.default_rng(11 + itype)
    for (h, w, cwin, ewin, lv) in [(12, 14, 5, 3, None), (13, 11, 3, 5, 4), (10, 16, 5, 7, 6)]:
        img = _img(rng, itype, h, w, lv)
        a = np.where(rng.random((h, w)) < 0.85, 255, 0).astype(np.uint8)
        kw = dict(w_exp=1.0, w_sat=0.2, w_con=0.5, w_ent=0.3, cwin=cwin, ewin=ewin)
        prm = enfuse_params(w_exposure=1.0, w_saturation=0.2, w_contrast=0.5, w_entropy=0.3,
                            contrast_window=cwin, entropy_window=ewin)
        got = oracle.enfuse_mask(img, a, itype, prm)
        want = direct_mask(img, a, itype, **kw)
        np.testing.assert_array_equal(got, want, err_msg=f"{w}x{h} c{cwin} e{ewin}")


def test_float_weights_structure(oracle):
    """Float images: exposure y = gray / FLT_MAX collapses to weight(0)
    (src/enfuse.h:637) and saturation uses max_value 1.0; contrast and
    entropy follow the integer structure (float sums, FLT_MAX scale)."""
    rng = np.random.default_rng(5)
    h, w = 12, 13
    img = (rng.random((h, w, 3)) * 0.4 + 0.05).astype(np.float32)
    a = np.full((h, w), 255, np.uint8)
    got = oracle.enfuse_mask(img, a, F32, enfuse_params(w_exposure=1.0))
    assert np.all(got == np.float32(math.exp(-0.5 * (0.5 / 0.2) ** 2)))
    got = oracle.enfuse_mask(img, a, F32, enfuse_params(w_exposure=0.0, w_saturation=1.0))
    mx, mn = img.max(2).astype(np.float64), img.min(2).astype(np.float64)
    np.testing.assert_array_equal(got, ((mx - mn) / (mx + mn)).astype(np.float32))
    # entropy of a window of distinct values is 1, scaled by FLT_MAX
    got = oracle.enfuse_mask(img, a, F32, enfuse_params(w_exposure=0.0, w_entropy=0.5, entropy_window=3))
    inner = got[1:-1, 1:-1]
    np.testing.assert_allclose(inner, 0.5 * np.finfo(np.float32).max, rtol=1e-6)
    assert np.all(got[0] == 0) and np.all(got[:, 0] == 0)


def test_fusion_identity_single_image(oracle):
    """K = 1 with zero weights: the normalised mask is the fallback 255/K =
    255, the mask coefficient is exactly 1 and collapse inverts the
    Laplacian, so the output is the input wherever alpha is set."""
    rng = np.random.default_rng(9)
    for itype in (U8, U16):
        img = _img(rng, itype, 37, 45)
        a = np.where(rng.random((37, 45)) < 0.8, 255, 0).astype(np.uint8)
        out, oa = oracle.enfuse([img], [a], itype, enfuse_params(w_exposure=0.0), oracle.auto_levels(45, 37))
        sel = a != 0
        np.testing.assert_array_equal(oa, a)
        np.testing.assert_array_equal(out[sel], img[sel])
        assert np.all(out[~sel] == 0)


def test_fusion_known_answer_constant_images(oracle):
    """Two constant u8 images with saturation weights 1/2 and 1/4 normalise
    to masks 170 and 85 (exact in float), i.e. coefficients 2/3 and 1/3 of
    the coarsest Gaussian level; every Laplacian level is 0."""
    h, w = 24, 40
    A = np.zeros((h, w, 3), np.uint8); A[...] = (90, 30, 30)       # saturation 1/2
    B = np.zeros((h, w, 3), np.uint8); B[...] = (150, 90, 90)      # saturation 1/4
    full = np.full((h, w), 255, np.uint8)
    prm = enfuse_params(w_exposure=0.0, w_saturation=1.0)
    out, _ = oracle.enfuse([A, B], [full, full], U8, prm, oracle.auto_levels(w, h), dither=0)
    cA = (170 << 7) / float(255 << 7)
    cB = (85 << 7) / float(255 << 7)
    for c in range(3):
        v = round_half_away(int(A[0, 0, c]) * 128 * cA) + round_half_away(int(B[0, 0, c]) * 128 * cB)
        q, f = v >> 7, v & 127
        want = q + 1 if f >= 64 else q
        assert np.all(out[..., c] == want), (c, out[0, 0], want)


def _noise_img(rng, itype, h, w, lo, hi):
    m = MAXV[itype]
    v = rng.uniform(lo, hi, size=(h, w, 3)) * m
    return np.clip(np.rint(v), 0, m).astype(np.uint8 if itype == U8 else np.uint16)


def test_hard_mask_dominant_image_passes_through(oracle):
    """Hard masks (src/enfuse.h:1402-1443): where one image has the largest
    weight everywhere it takes every pixel whole (mask 255, coefficient
    exactly 1; the others 0), so the fused output is that image; the soft
    masks of the same stack blend."""
    rng = np.random.default_rng(21)
    h, w = 33, 47
    for itype in (U8, U16):
        dark = _noise_img(rng, itype, h, w, 0.0, 0.05)
        mid = _noise_img(rng, itype, h, w, 0.45, 0.55)
        bright = _noise_img(rng, itype, h, w, 0.95, 1.0)
        full = [np.full((h, w), 255, np.uint8)] * 3
        L = oracle.auto_levels(w, h)
        out, _ = oracle.enfuse([dark, mid, bright], full, itype, enfuse_params(w_exposure=1.0, hard_mask=True), L,
                               dither=1, seed=3)
        np.testing.assert_array_equal(out, mid)
        soft, _ = oracle.enfuse([dark, mid, bright], full, itype, enfuse_params(w_exposure=1.0), L, dither=1, seed=3)
        assert not np.array_equal(soft, mid)


def test_hard_mask_ties_go_to_the_first_image(oracle):
    """Equal weights: the first image of largest weight wins (strict >).
    B is A with its channels rotated, so gray and saturation, hence every
    weight, are equal, and each order of the stack returns its first image."""
    rng = np.random.default_rng(22)
    h, w = 29, 31
    A = _noise_img(rng, U8, h, w, 0.1, 0.9)
    B = np.ascontiguousarray(np.roll(A, 1, axis=2))
    full = [np.full((h, w), 255, np.uint8)] * 2
    prm = enfuse_params(w_exposure=1.0, w_saturation=0.2, hard_mask=True)
    L = oracle.auto_levels(w, h)
    np.testing.assert_array_equal(oracle.enfuse([A, B], full, U8, prm, L)[0], A)
    np.testing.assert_array_equal(oracle.enfuse([B, A], full, U8, prm, L)[0], B)


def test_hard_mask_zero_weights_take_float_share(oracle):
    """No weight above 0: every image gets float(255) / K, a float division
    (src/enfuse.h:1430-1432), where the soft fallback is the integer 255 / K
    (:1534).  Seven constant images of value 100 then fuse to 100 with hard
    masks (36.428574 -> 4663/32640 per image) and to 99 with soft masks
    (36 -> 4608/32640), computed here in closed form (no dither)."""
    h, w, K, v = 20, 24, 7, 100
    imgs = [np.full((h, w, 3), v, np.uint8) for _ in range(K)]
    full = [np.full((h, w), 255, np.uint8)] * K
    L = oracle.auto_levels(w, h)

    def closed_form(mask_value):
        mfix = round_half_away(float(mask_value) * 128.0)
        c = mfix / float(255 << 7)
        s = K * round_half_away(v * 128 * c)
        q, f = s >> 7, s & 127
        return q + 1 if f >= 64 else q

    want_hard = closed_form(np.float32(255) / np.float32(K))
    want_soft = closed_form(255 // K)
    assert (want_hard, want_soft) == (100, 99)
    hard, _ = oracle.enfuse(imgs, full, U8, enfuse_params(w_exposure=0.0, hard_mask=True), L, dither=0)
    soft, _ = oracle.enfuse(imgs, full, U8, enfuse_params(w_exposure=0.0), L, dither=0)
    assert np.all(hard == want_hard) and np.all(soft == want_soft)


# ---- float local standard deviation (src/enfuse.h:75-219 with
# RealPromote<float> = float): pinned by closed forms where every partial
# sum is exact, and by known answers that only the reference's float running
# sums in its sliding update order produce.

def _ulp32(x):
    x = np.abs(np.asarray(x, np.float32))
    return (np.nextafter(x, np.float32(np.inf)) - x).astype(np.float64)


def _closed_form_sd(g, m, n):
    """Exact sample standard deviation of the masked window samples, the
    window sums, count and the border/mask rule (interior pixels under the
    mask only; fewer than two samples give 0)."""
    h, w = g.shape
    b = n // 2
    sd = np.zeros((h, w))
    sums = np.zeros((h, w))
    sqs = np.zeros((h, w))
    cnt = np.zeros((h, w), np.int64)
    written = np.zeros((h, w), bool)
    for y in range(b, h - b):
        for x in range(b, w - b):
            if not m[y, x]:
                continue
            win = g[y - b:y + b + 1, x - b:x + b + 1].astype(np.float64)
            sel = win[m[y - b:y + b + 1, x - b:x + b + 1] != 0]
            written[y, x] = True
            cnt[y, x] = sel.size
            sums[y, x], sqs[y, x] = sel.sum(), (sel * sel).sum()
            if sel.size > 1:
                sd[y, x] = math.sqrt(max(0.0, (sqs[y, x] - sums[y, x] ** 2 / sel.size) / (sel.size - 1)))
    return sd, sums, sqs, cnt, written


@pytest.mark.parametrize("n", [3, 5, 7])
@pytest.mark.parametrize("planted", [0, 1])
def test_float_stddev_exact_sums_match_closed_form(oracle, n, planted):
    """Samples k/8 (k <= 16): every column sum, running sum and sum of
    squares is exact in float, so only the final formula rounds; the result
    must equal the exact sample standard deviation within those roundings
    (RN(sum^2 / n), the subtraction, the division by n - 1 and the square
    root), pixels with fewer than two samples give exactly 0, borders and
    pixels outside the mask stay 0.  A double accumulator passes this test
    too (the sums are exact); the known answers below tell them apart."""
    rng = np.random.default_rng(40 + n)
    h, w = 23, 31
    g = (rng.integers(0, 17, size=(h, w)) / 8.0).astype(np.float32)
    m = (rng.random((h, w)) < 0.8).astype(np.uint8)
    m[5, 5:5 + n] = 0                               # sparse windows: counts 0..n
    m[6:6 + n, 9] = 0
    m[10:10 + n, 12:12 + n] = 0
    m[10 + n // 2, 12 + n // 2] = 1                 # a window with one sample
    got = oracle.local_stddev_f32(g, m, n, planted).astype(np.float64)
    sd, sums, sqs, cnt, written = _closed_form_sd(g, m, n)
    assert np.all(got[~written] == 0)
    assert np.all(got[written & (cnt <= 1)] == 0)
    assert np.any(written & (cnt == 1))
    sel = written & (cnt > 1)
    c = cnt[sel].astype(np.float64)
    # rounding of sum^2 / n and of the subtraction, carried through / (n-1)
    tol_var = (_ulp32(sums[sel] ** 2 / c) + _ulp32(sqs[sel])) / (c - 1) + _ulp32(sd[sel] ** 2)
    err_var = np.abs(got[sel] ** 2 - sd[sel] ** 2)
    assert np.all(err_var <= tol_var + 2 * got[sel] * _ulp32(got[sel])), (err_var - tol_var).max()


def test_float_stddev_known_answer_float_sums():
    """A 5x5 window of 4097.0f (n = 5), by hand in float: 4097^2 =
    16785409 rounds to 16785408; the column sums of squares grow exactly to
    83927040, the window's to 419635200 (sum 102425 exact).  sum^2 =
    10490880625 rounds to 10490881024, / 25 to 419635232, and
    419635200 - 419635232 = -32, so the variance is negative and the result
    NaN.  Double sums give exactly 0 (the window is constant)."""
    from oracle_binding import load_oracle
    orc = load_oracle()
    g = np.full((5, 5), 4097.0, np.float32)
    m = np.ones((5, 5), np.uint8)
    got = orc.local_stddev_f32(g, m, 5)
    assert np.isnan(got[2, 2])
    assert np.all(np.delete(got.ravel(), 12) == 0)
    # the planted double accumulator misses it
    assert orc.local_stddev_f32(g, m, 5, planted=1)[2, 2] == 0.0


def test_float_stddev_known_answer_sliding_update():
    """5 x 6 image, n = 5, column 0 = 8192.0f, columns 1-5 = 1.0f; outputs
    at (2, 2) and (2, 3).  First window: sum = 5 * 8192 + 20 = 40980 (exact),
    sum of squares 5 * 2^26 + 20 rounds to 335544320 (spacing 32), sum^2 =
    1679360400 rounds to 1679360384, / 25 to 67174416, the difference
    268369904 / 24 to 11182079: result RN(sqrt(11182079)).  Slide by one
    column in the reference's order, sum += s_new - s_old: the squares'
    update 5 - 335544320 rounds to -335544320, so the running sum of
    squares becomes 0 while the sum becomes exactly 25: (0 - 625 / 25) / 24
    < 0 and the result is NaN.  Summing the window anew (or in double)
    gives 25 - 25 = 0, i.e. 0."""
    from oracle_binding import load_oracle
    orc = load_oracle()
    g = np.ones((5, 6), np.float32)
    g[:, 0] = 8192.0
    m = np.ones((5, 6), np.uint8)
    got = orc.local_stddev_f32(g, m, 5)
    assert got[2, 2] == np.sqrt(np.float32(11182079.0), dtype=np.float32)
    assert np.isnan(got[2, 3])
    for planted in (1, 2):
        assert orc.local_stddev_f32(g, m, 5, planted=planted)[2, 3] == 0.0


def test_float_contrast_weight_is_the_float_stddev(oracle):
    """The contrast weight of a float image with equal channels is
    float(w_con * sd) of its own gray values (average projector: (3v)/3 =
    v exactly for values whose triple is exact in float), through
    enfuse_mask, on the known answers and a random image."""
    rng = np.random.default_rng(3)
    for g in (np.full((5, 5), 4097.0, np.float32), np.where(np.arange(6) == 0, 8192.0, 1.0).astype(np.float32)[None, :]
              .repeat(5, 0), (rng.integers(0, 1 << 12, size=(17, 19)) / 64.0).astype(np.float32)):
        h, w = g.shape
        m = np.full((h, w), 255, np.uint8)
        img = np.repeat(g[..., None], 3, axis=2)
        got = oracle.enfuse_mask(img, m, F32, enfuse_params(w_exposure=0.0, w_contrast=1.0, contrast_window=5))
        want = oracle.local_stddev_f32(g, m, 5)
        np.testing.assert_array_equal(got, want)
