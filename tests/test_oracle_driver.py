"""Pins the oracle's restatement of enblend's driver around the blend (import
thresholds, assembly, overlap inspection, seam-mask bounds, the blending-loop
step and the checkpoint conversion) with closed forms, brute force in numpy on
small canvases, and the reference's own special cases."""
import numpy as np
import pytest

from oracle_binding import F32, U8, U16, U32, load_oracle

NP = {U8: np.uint8, U16: np.uint16, F32: np.float32}
FMAX = {U8: 255, U16: 65535, F32: 1.0}


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


def _samples(rng, h, w, bands, ftype):
    if ftype == F32:
        return rng.uniform(0.0, 1.0, size=(h, w, bands)).astype(np.float32)
    return rng.integers(0, FMAX[ftype] + 1, size=(h, w, bands)).astype(NP[ftype])


def _layer(rng, h, w, ftype, C=3, alpha="blob"):
    a = _samples(rng, h, w, C + (alpha is not None), ftype)
    if alpha == "blob":
        yy, xx = np.mgrid[0:h, 0:w]
        inside = (xx - w / 2) ** 2 / (w / 2.2) ** 2 + (yy - h / 2) ** 2 / (h / 2.2) ** 2 <= 1
        a[..., C] = np.where(inside, FMAX[ftype], 0).astype(a.dtype)
    return a


# ---- import: thresholds and range mapping (src/assemble.h:231-289) ----------

@pytest.mark.parametrize("ftype,itype,cases", [
    # Threshold(T(max/2), T(max)): integer image types truncate max/2
    (U8, U8, [(0, 0), (126, 0), (127, 255), (200, 255), (255, 255)]),
    (U16, U16, [(0, 0), (32766, 0), (32767, 255), (65535, 255)]),
    (U8, U16, [(126, 0), (127, 255), (255, 255)]),
    # a float image keeps max/2 = 127.5: 127 is below it
    (U8, F32, [(127, 0), (128, 255), (255, 255)]),
    (U16, F32, [(32767, 0), (32768, 255)]),
    (F32, F32, [(0.4999999, 0), (0.5, 255), (1.0, 255), (1.0000001, 0), (float("nan"), 255), (-0.5, 0)]),
])
def test_import_alpha_threshold(orc, ftype, itype, cases):
    vals = np.array([v for v, _ in cases], dtype=NP[ftype])
    lay = np.zeros((1, len(cases), 4), NP[ftype])
    lay[0, :, 3] = vals
    n, img, al, bb, _ = orc.assemble([(lay, 0, 0, ftype)], itype, 3, len(cases), 1)
    assert n == 1
    assert al[0].tolist() == [e for _, e in cases]


def test_import_without_alpha_sets_every_pixel(orc):
    rng = np.random.default_rng(0)
    lay = _samples(rng, 5, 7, 3, U8)
    n, img, al, bb, _ = orc.assemble([(lay, 2, 1, U8)], U8, 3, 12, 9)
    assert (al[1:6, 2:9] == 255).all() and al.sum() == 255 * 35
    assert bb == (2, 1, 9, 6)
    np.testing.assert_array_equal(img[1:6, 2:9], lay)


def test_import_range_mapping_closed_forms(orc):
    x = np.arange(256, dtype=np.uint8)
    lay = np.zeros((1, 256, 3), np.uint8)
    lay[0, :, 0] = x
    lay[0, :, 1] = x[::-1]
    lay[0, :, 2] = 7
    # 8-bit file into a 16-bit image: linearRangeMapping scale 257, offset 0
    _, img16, _, _, _ = orc.assemble([(lay, 0, 0, U8)], U16, 3, 256, 1)
    np.testing.assert_array_equal(img16[0, :, 0], x.astype(np.uint32) * 257)
    np.testing.assert_array_equal(img16[0, :, 1], x[::-1].astype(np.uint32) * 257)
    # into float: float multiplier 1/255 computed in float, one float product
    _, imgf, _, _, _ = orc.assemble([(lay, 0, 0, U8)], F32, 3, 256, 1)
    scale = np.float32(1.0) / np.float32(255.0)
    np.testing.assert_array_equal(imgf[0, :, 0], (scale * x.astype(np.float32)).astype(np.float32))
    assert imgf[0, 255, 0] == np.float32(scale * np.float32(255))
    # 16-bit file into float
    y = np.array([0, 1, 2, 32767, 65534, 65535], np.uint16)
    lay16 = np.zeros((1, len(y), 3), np.uint16)
    lay16[0, :, 0] = y
    _, imgf, _, _, _ = orc.assemble([(lay16, 0, 0, U16)], F32, 3, len(y), 1)
    s16 = np.float32(1.0) / np.float32(65535.0)
    np.testing.assert_array_equal(imgf[0, :, 0], (s16 * y.astype(np.float32)).astype(np.float32))


def test_unsupported_layer_types(orc):
    lay = np.zeros((2, 2, 4), np.float32)
    with pytest.raises(ValueError):
        orc.assemble([(lay, 0, 0, F32)], U16, 3, 2, 2)
    lay = np.zeros((2, 2, 4), np.uint16)
    with pytest.raises(ValueError):
        orc.assemble([(lay, 0, 0, U16)], U8, 3, 2, 2)


# ---- assemble: greedy non-overlapping pre-assembly (src/assemble.h:300-444) --

def _assemble_numpy(layers, itype, C, uw, uh, taken, one_at_a_time):
    """Brute force over numpy arrays: thresholded alphas, greedy acceptance."""
    img = np.zeros((uh, uw, C), NP[itype])
    al = np.zeros((uh, uw), np.uint8)
    todo = [i for i in range(len(layers)) if not taken[i]]
    if not todo:
        return 0, None, None, None, taken
    taken = list(taken)
    placed = 0
    for k, i in enumerate(todo):
        a, x, y, ft = layers[i]
        h, w, bands = a.shape
        im = a[..., :C].astype(NP[itype])
        if bands > C:
            t = NP[itype]
            lo, hi = t(FMAX[ft] / 2.0), t(FMAX[ft])
            v = a[..., C].astype(t)
            m = ~((v < lo) | (hi < v))
        else:
            m = np.ones((h, w), bool)
        if k == 0:
            img[y:y + h, x:x + w] = im
            al[y:y + h, x:x + w] = np.where(m, 255, 0)
        else:
            if one_at_a_time:
                break
            if (m & (al[y:y + h, x:x + w] != 0)).any():
                continue
            img[y:y + h, x:x + w][m] = im[m]
            al[y:y + h, x:x + w][m] = 255
        taken[i] = 1
        placed += 1
    ys, xs = np.nonzero(al)
    bb = (0, 0, 0, 0) if len(xs) == 0 else (int(xs.min()), int(ys.min()), int(xs.max()) + 1, int(ys.max()) + 1)
    return placed, img, al, bb, taken


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("one_at_a_time", [False, True])
def test_assemble_matches_brute_force(orc, itype, one_at_a_time):
    rng = np.random.default_rng(11 + itype)
    uw, uh = 61, 37
    specs = [(20, 25, 0, 0), (18, 22, 30, 3), (15, 15, 10, 10), (17, 20, 40, 0), (9, 8, 52, 28), (12, 12, 1, 24)]
    layers = [(_layer(rng, h, w, itype), x, y, itype) for h, w, x, y in specs]
    taken = [0] * len(layers)
    calls = 0
    while not all(taken):
        n, img, al, bb, tk = orc.assemble(layers, itype, 3, uw, uh, taken, one_at_a_time)
        en, eimg, eal, ebb, etk = _assemble_numpy(layers, itype, 3, uw, uh, taken, one_at_a_time)
        assert n == en and bb == ebb and list(tk) == list(etk)
        np.testing.assert_array_equal(al, eal)
        np.testing.assert_array_equal(img, eimg)
        taken = list(tk)
        calls += 1
    assert calls == (len(layers) if one_at_a_time else calls)
    n, *_ = orc.assemble(layers, itype, 3, uw, uh, taken, one_at_a_time)
    assert n == 0


def test_assemble_greedy_order(orc):
    """Layer 1 overlaps layer 0, layer 2 overlaps neither: the first call
    takes 0 and 2, the second takes 1 alone."""
    rng = np.random.default_rng(3)
    a0 = _layer(rng, 10, 10, U8, alpha=None)
    a1 = _layer(rng, 10, 10, U8, alpha=None)
    a2 = _layer(rng, 10, 10, U8, alpha=None)
    layers = [(a0, 0, 0, U8), (a1, 5, 5, U8), (a2, 20, 0, U8)]
    n, img, al, bb, tk = orc.assemble(layers, U8, 3, 30, 15)
    assert n == 2 and list(tk) == [1, 0, 1] and bb == (0, 0, 30, 10)
    n, img, al, bb, tk = orc.assemble(layers, U8, 3, 30, 15, taken=tk)
    assert n == 1 and list(tk) == [1, 1, 1] and bb == (5, 5, 15, 15)


def test_assemble_first_layer_writes_its_whole_rectangle(orc):
    """The first layer is imported straight onto the canvas: its samples under
    a zero alpha are kept; later layers go through copyImageIf."""
    rng = np.random.default_rng(4)
    a0 = _layer(rng, 8, 8, U16)
    a1 = _layer(rng, 8, 8, U16)
    layers = [(a0, 0, 0, U16), (a1, 12, 0, U16)]
    n, img, al, bb, tk = orc.assemble(layers, U16, 3, 20, 8)
    assert n == 2
    np.testing.assert_array_equal(img[:, :8], a0[..., :3])
    m1 = a1[..., 3] != 0
    np.testing.assert_array_equal(img[:, 12:][m1], a1[..., :3][m1])
    assert (img[:, 12:][~m1] == 0).all()


# ---- inspectOverlap and maskBounds -------------------------------------------

def test_inspect_overlap_cases(orc):
    rng = np.random.default_rng(5)
    for _ in range(50):
        a1 = (rng.random((9, 11)) < rng.random()).astype(np.uint8) * 255
        a2 = (rng.random((9, 11)) < rng.random()).astype(np.uint8) * 255
        both = ((a1 != 0) & (a2 != 0)).any()
        only2 = ((a2 != 0) & (a1 == 0)).any()
        expect = 1 if both and only2 else 2 if both else 0
        assert orc.inspect_overlap(a1, a2) == expect
    z = np.zeros((4, 4), np.uint8)
    f = np.full((4, 4), 255, np.uint8)
    assert orc.inspect_overlap(f, f) == 2
    assert orc.inspect_overlap(f, z) == 0
    assert orc.inspect_overlap(z, f) == 0


def _mask_bounds_numpy(m):
    h, w = m.shape
    boxes = []
    ys, xs = np.nonzero(m[:, 1:] != m[:, :-1])
    for y, x in zip(ys, xs + 1):
        boxes.append((x - 1, y, x + 1, y + 1))
    ys, xs = np.nonzero(m[1:, :] != m[:-1, :])
    for y, x in zip(ys + 1, xs):
        boxes.append((x, y - 1, x + 1, y + 1))
    if not boxes:
        return False, None
    b = np.array(boxes)
    return True, (int(b[:, 0].min()), int(b[:, 1].min()), int(b[:, 2].max()), int(b[:, 3].max()))


def test_mask_bounds(orc):
    rng = np.random.default_rng(6)
    for _ in range(40):
        h, w = rng.integers(1, 20, size=2)
        m = np.zeros((h, w), np.uint8)
        if rng.random() < 0.8:
            x0, x1 = sorted(rng.integers(0, w + 1, size=2))
            y0, y1 = sorted(rng.integers(0, h + 1, size=2))
            m[y0:y1, x0:x1] = 255
        ok, bb = orc.mask_bounds(m)
        eok, ebb = _mask_bounds_numpy(m)
        assert ok == eok
        if ok:
            assert bb == ebb
    # a vertical seam at column k: transitions between k-1 and k on every row
    m = np.zeros((6, 10), np.uint8)
    m[:, 4:] = 255
    assert orc.mask_bounds(m) == (True, (3, 0, 5, 6))


# ---- the blending-loop step (src/enblend.h:127-545) -------------------------

def _pair_scene(rng, itype, uw=90, uh=40, ov=(30, 60)):
    C = 3
    black = np.zeros((uh, uw, C), NP[itype])
    white = np.zeros((uh, uw, C), NP[itype])
    ba = np.zeros((uh, uw), np.uint8)
    wa = np.zeros((uh, uw), np.uint8)
    black[:, :ov[1]] = _samples(rng, uh, ov[1], C, itype)
    white[:, ov[0]:] = _samples(rng, uh, uw - ov[0], C, itype)
    ba[2:uh - 3, 1:ov[1]] = 255
    wa[4:uh - 1, ov[0]:uw - 2] = 255
    return black, ba, white, wa


def _bbox(a):
    ys, xs = np.nonzero(a)
    return (int(xs.min()), int(ys.min()), int(xs.max()) + 1, int(ys.max()) + 1)


@pytest.mark.parametrize("itype", [U8, U16, F32])
def test_step_blends_on_the_union_box(orc, itype):
    """The step equals the NFT mask of the union-box alphas followed by the
    pair step on that box (roiBounds' border covers the whole union box)."""
    rng = np.random.default_rng(7 + itype)
    black, ba, white, wa = _pair_scene(rng, itype)
    bbb, wbb = _bbox(ba), _bbox(wa)
    rc, outcome, b, bal, nbb, info = orc.enblend_step(black, ba, bbb, white, wa, wbb, itype, seed=5)
    assert rc == 0 and outcome == 0
    u = (min(bbb[0], wbb[0]), min(bbb[1], wbb[1]), max(bbb[2], wbb[2]), max(bbb[3], wbb[3]))
    assert nbb == u and tuple(info[4:8]) == u
    x0, y0, x1, y1 = u
    mask, mrc = orc.nft_mask(wa[y0:y1, x0:x1], ba[y0:y1, x0:x1])
    assert mrc == 0
    ok, mb = orc.mask_bounds(mask)
    assert ok and tuple(info[:4]) == (mb[0] + x0, mb[1] + y0, mb[2] + x0, mb[3] + y0)
    levels = orc.auto_levels(x1 - x0, y1 - y0)
    assert info[8] == levels
    eb, eba = orc.enblend_pair(white[y0:y1, x0:x1], wa[y0:y1, x0:x1], black[y0:y1, x0:x1], ba[y0:y1, x0:x1],
                               mask, levels, False, itype, 1, 5)
    exp = black.copy()
    expa = ba.copy()
    exp[y0:y1, x0:x1] = eb
    expa[y0:y1, x0:x1] = eba
    np.testing.assert_array_equal(b, exp)
    np.testing.assert_array_equal(bal, expa)


def test_step_with_wraparound_uses_the_strip_mask(orc):
    """A union box as wide as the canvas with wraparound: the horizontal-strip
    NFT mask and the pair step with wraparound on the union box."""
    rng = np.random.default_rng(21)
    uw, uh = 96, 40
    black = _samples(rng, uh, uw, 3, U16)
    white = _samples(rng, uh, uw, 3, U16)
    ba = np.zeros((uh, uw), np.uint8)
    wa = np.zeros((uh, uw), np.uint8)
    ba[1:uh - 1, : int(0.62 * uw)] = 255
    wa[2:uh - 2, int(0.40 * uw):] = 255
    bbb, wbb = _bbox(ba), _bbox(wa)
    rc, outcome, b, bal, nbb, info = orc.enblend_step(black, ba, bbb, white, wa, wbb, U16, wrap=True, seed=2)
    assert rc == 0 and outcome == 0 and nbb[0] == 0 and nbb[2] == uw
    x0, y0, x1, y1 = nbb
    mask, mrc = orc.nft_mask_periodic(wa[y0:y1, x0:x1], ba[y0:y1, x0:x1], 1)
    assert mrc == 0
    assert (mask != orc.nft_mask(wa[y0:y1, x0:x1], ba[y0:y1, x0:x1])[0]).any()
    levels = orc.auto_levels(x1 - x0, y1 - y0)
    eb, eba = orc.enblend_pair(white[y0:y1], wa[y0:y1], black[y0:y1], ba[y0:y1], mask, levels, True, U16, 1, 2)
    exp, expa = black.copy(), ba.copy()
    exp[y0:y1], expa[y0:y1] = eb, eba
    np.testing.assert_array_equal(b, exp)
    np.testing.assert_array_equal(bal, expa)


def test_step_redundant_and_disjoint(orc):
    rng = np.random.default_rng(8)
    black, ba, white, wa = _pair_scene(rng, U8)
    # white inside black: CompleteOverlap, black untouched
    wa2 = np.zeros_like(wa)
    wa2[10:20, 10:20] = 255
    rc, outcome, b, bal, nbb, _ = orc.enblend_step(black, ba, _bbox(ba), white, wa2, _bbox(wa2), U8)
    assert rc == 0 and outcome == 1 and nbb == _bbox(ba)
    np.testing.assert_array_equal(b, black)
    np.testing.assert_array_equal(bal, ba)
    # disjoint: white copied where its alpha is set, alpha union, box union
    wa3 = np.zeros_like(wa)
    wa3[5:15, 70:85] = 255
    rc, outcome, b, bal, nbb, _ = orc.enblend_step(black, ba, _bbox(ba), white, wa3, _bbox(wa3), U8)
    assert rc == 0 and outcome == 2
    m = wa3 != 0
    exp = black.copy()
    exp[m] = white[m]
    np.testing.assert_array_equal(b, exp)
    np.testing.assert_array_equal(bal, np.maximum(ba, wa3))
    assert nbb == (1, 2, 85, 37)
    # disjoint with --levels: blended instead of copied
    rc, outcome, *_ = orc.enblend_step(black, ba, _bbox(ba), white, wa3, _bbox(wa3), U8, exact_levels=3)
    assert rc == 0 and outcome == 0


# ---- checkpoint (src/assemble.h:130-224) -------------------------------------

def test_checkpoint_identity_and_alpha(orc):
    rng = np.random.default_rng(9)
    for it in (U8, U16, F32):
        img = _samples(rng, 5, 6, 3, it)
        al = (rng.random((5, 6)) < 0.5).astype(np.uint8) * rng.integers(1, 256, size=(5, 6)).astype(np.uint8)
        out, ct = orc.checkpoint(img, al, it, it)
        assert ct == it
        np.testing.assert_array_equal(out[..., :3], img)
        amax = {U8: 255, U16: 65535, F32: 1.0}[it]
        np.testing.assert_array_equal(out[..., 3], np.where(al != 0, amax, 0).astype(out.dtype))


def test_checkpoint_narrowing_closed_forms(orc):
    x = np.arange(65536, dtype=np.uint16).reshape(256, 256, 1).repeat(3, axis=2)
    al = np.ones((256, 256), np.uint8)
    out, ct = orc.checkpoint(x, al, U16, U8)
    # linearRangeMapping(0, 65535 -> 0, 255) in UInt16: round(x / 257) exactly
    assert ct == U16
    np.testing.assert_array_equal(out[..., 0], (x[..., 0].astype(np.int64) + 128) // 257)
    assert out[..., 3].max() == 65535
    # 8-bit image exported as float: narrowed to {0, 1} in the 8-bit type
    y = np.arange(256, dtype=np.uint8).reshape(16, 16, 1).repeat(3, axis=2)
    out, ct = orc.checkpoint(y, np.ones((16, 16), np.uint8), U8, F32)
    assert ct == U8
    np.testing.assert_array_equal(out[..., 0], (y[..., 0] >= 128).astype(np.uint8))


def test_checkpoint_float_rescale(orc):
    v = np.array([0.0, 0.25, 0.5, 1.0, 1.0 / 255, 2.0, -0.5, np.inf, -np.inf, np.nan, 0.001960784], np.float32)
    img = np.repeat(v.reshape(1, -1, 1), 3, axis=2)
    al = np.ones((1, len(v)), np.uint8)
    out, ct = orc.checkpoint(img, al, F32, U8)
    assert ct == U32
    d = v.astype(np.float64) * 255.0
    exp = np.where(np.isnan(d) | (d < 0), 0, np.where(d > 4294967295.0, 4294967295.0, np.floor(d + 0.5)))
    np.testing.assert_array_equal(out[0, :, 0], exp.astype(np.uint32))
    assert out[0, 0, 3] == 4294967295
    out, ct = orc.checkpoint(img, al, F32, U16)
    d = v.astype(np.float64) * 65535.0
    exp = np.where(np.isnan(d) | (d < 0), 0, np.where(d > 4294967295.0, 4294967295.0, np.floor(d + 0.5)))
    np.testing.assert_array_equal(out[0, :, 0], exp.astype(np.uint32))


def test_checkpoint_unsupported(orc):
    with pytest.raises(ValueError):
        orc.checkpoint(np.zeros((2, 2, 3), np.uint8), np.zeros((2, 2), np.uint8), U8, U16)
