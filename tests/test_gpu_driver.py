"""GPU driver around the blend (driver.hip) against the oracle, bit for bit:
layer import and pre-assembly, overlap inspection, seam-mask bounds, one
blending-loop step, the whole enblend loop over several layers, and the
checkpoint conversion."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle_binding import F32, U8, U16, U32  # noqa: E402

NP = {U8: np.uint8, U16: np.uint16, F32: np.float32, U32: np.uint32}
FMAX = {U8: 255, U16: 65535, F32: 1.0}


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _samples(rng, h, w, bands, ftype):
    if ftype == F32:
        return rng.uniform(0.0, 1.0, size=(h, w, bands)).astype(np.float32)
    return rng.integers(0, FMAX[ftype] + 1, size=(h, w, bands)).astype(NP[ftype])


def _layer(rng, h, w, ftype, C=3, alpha=True):
    a = _samples(rng, h, w, C + bool(alpha), ftype)
    if alpha:
        yy, xx = np.mgrid[0:h, 0:w]
        r = (xx - w / 2) ** 2 / (w / 2.2) ** 2 + (yy - h / 2) ** 2 / (h / 2.2) ** 2
        # threshold neighbourhood values on the rim, full alpha inside
        if ftype == F32:
            a[..., C] = np.where(r <= 1, 1.0, np.where(r <= 1.2, 0.5, 0.49)).astype(np.float32)
        else:
            half = FMAX[ftype] // 2
            a[..., C] = np.where(r <= 1, FMAX[ftype], np.where(r <= 1.2, half, half - 1)).astype(a.dtype)
    return a


def _gpu_assemble(enbl, ctx, layers, itype, C, uw, uh, taken, one_at_a_time):
    dl = [(_dev(a), x, y, ft) for a, x, y, ft in layers]
    tab = enbl.layer_table(dl)
    tk = (ctypes.c_int * len(layers))(*taken)
    # one-at-a-time runs use 16-byte aligned alpha rows (the wide bounding-box
    # pass), the others odd pitches
    ipitch, apitch = uw * C + 5, ((uw + 15) // 16 + 1) * 16 if one_at_a_time else uw + 3
    img = torch.full((uh, ipitch), 77, dtype=getattr(torch, np.dtype(NP[itype]).name), device="cuda")
    al = torch.full((uh, apitch), 9, dtype=torch.uint8, device="cuda")
    bb = enbl.Rect()
    torch.cuda.synchronize()
    n = enbl.lib().enbl_assemble(ctx.h, itype, C, uw, uh, len(layers), tab, int(one_at_a_time), tk, _vp(img), ipitch,
                                 _vp(al), apitch, ctypes.byref(bb))
    if n < 0:
        enbl.check(n)
    ctx.sync()
    img = img.cpu().numpy()
    al = al.cpu().numpy()
    if n > 0:
        assert (img[:, uw * C:] == 77).all() and (al[:, uw:] == 9).all(), "wrote past the rows"
    return n, img[:, :uw * C].reshape(uh, uw, C), al[:, :uw], bb.tuple(), list(tk)


@pytest.mark.parametrize("itype,ftypes", [(U8, [U8]), (U16, [U16, U8]), (F32, [F32, U8, U16])])
@pytest.mark.parametrize("C", [3, 1])
@pytest.mark.parametrize("one_at_a_time", [False, True])
def test_assemble_matches_oracle(enbl, gpu_ctx, oracle, itype, ftypes, C, one_at_a_time):
    print(f"assemble itype={itype} files={ftypes} C={C} one_at_a_time={one_at_a_time}", flush=True)
    rng = np.random.default_rng(100 + itype * 3 + C + 7 * one_at_a_time)
    uw, uh = 301, 133
    specs = [(60, 90, 3, 5), (70, 80, 81, 41), (40, 50, 160, 1), (50, 70, 230, 70), (30, 31, 20, 100),
             (25, 33, 265, 2), (20, 20, 120, 110)]
    layers = []
    for k, (h, w, x, y) in enumerate(specs):
        ft = ftypes[k % len(ftypes)]
        layers.append((_layer(rng, h, w, ft, C=C, alpha=(k != 4)), x, y, ft))
    taken = [0] * len(layers)
    for _ in range(len(layers) + 1):
        en, eimg, eal, ebb, etk = oracle.assemble(layers, itype, C, uw, uh, taken, one_at_a_time)
        n, img, al, bb, tk = _gpu_assemble(enbl, gpu_ctx, layers, itype, C, uw, uh, taken, one_at_a_time)
        assert n == en and tk == list(etk)
        if n == 0:
            break
        assert bb == ebb
        np.testing.assert_array_equal(al, eal)
        np.testing.assert_array_equal(img, eimg)
        taken = tk
    assert all(taken)


def test_assemble_rejects_bad_layers(enbl, gpu_ctx):
    lay = _dev(np.zeros((4, 4, 4), np.float32))
    tab = enbl.layer_table([(lay, 0, 0, F32)])
    tk = (ctypes.c_int * 1)(0)
    img = torch.zeros((4, 12), dtype=torch.uint16, device="cuda")
    al = torch.zeros((4, 4), dtype=torch.uint8, device="cuda")
    bb = enbl.Rect()
    rc = enbl.lib().enbl_assemble(gpu_ctx.h, U16, 3, 4, 4, 1, tab, 0, tk, _vp(img), 12, _vp(al), 4, ctypes.byref(bb))
    assert rc == enbl.ENBL_E_UNSUPPORTED
    tab = enbl.layer_table([(_dev(np.zeros((4, 4, 4), np.uint16)), 1, 0, U16)])
    rc = enbl.lib().enbl_assemble(gpu_ctx.h, U16, 3, 4, 4, 1, tab, 0, tk, _vp(img), 12, _vp(al), 4, ctypes.byref(bb))
    assert rc == enbl.ENBL_E_ARG


@pytest.mark.parametrize("shape", [(1, 1), (7, 300), (300, 7), (257, 513), (1000, 1100)])
def test_inspect_overlap_and_mask_bounds(enbl, gpu_ctx, oracle, shape):
    h, w = shape
    print(f"overlap / mask bounds {w}x{h}", flush=True)
    rng = np.random.default_rng(h * 31 + w)
    L = enbl.lib()
    for trial in range(6):
        a1 = np.zeros((h, w), np.uint8)
        a2 = np.zeros((h, w), np.uint8)
        for a in (a1, a2):
            if trial % 3:
                y0, y1 = sorted(rng.integers(0, h + 1, size=2))
                x0, x1 = sorted(rng.integers(0, w + 1, size=2))
                a[y0:y1, x0:x1] = rng.integers(1, 256)
        if trial == 5:
            a2[:] = a1
        d1, d2 = _dev(a1), _dev(a2)
        ov = ctypes.c_int(-1)
        enbl.check(L.enbl_inspect_overlap(gpu_ctx.h, w, h, _vp(d1), w, _vp(d2), w, ctypes.byref(ov)))
        assert ov.value == oracle.inspect_overlap(a1, a2)
        bb = enbl.Rect()
        found = ctypes.c_int(-1)
        enbl.check(L.enbl_mask_bounds(gpu_ctx.h, w, h, _vp(d1), w, ctypes.byref(bb), ctypes.byref(found)))
        ok, ebb = oracle.mask_bounds(a1)
        assert bool(found.value) == ok
        if ok:
            assert bb.tuple() == ebb


def _pair_scene(rng, itype, uw, uh, bx, wx, C=3):
    """Black on columns [0, bx), white on [wx, uw), elliptic alphas with
    their centres apart so the NFT has exclusive pixels on both sides."""
    black = np.zeros((uh, uw, C), NP[itype])
    white = np.zeros((uh, uw, C), NP[itype])
    black[:, :bx] = _samples(rng, uh, bx, C, itype)
    white[:, wx:] = _samples(rng, uh, uw - wx, C, itype)
    yy, xx = np.mgrid[0:uh, 0:uw]
    ba = (((xx - bx / 2) / (bx / 2.05)) ** 2 + ((yy - uh / 2) / (uh / 2.1)) ** 2 <= 1).astype(np.uint8) * 255
    cw = (wx + uw) / 2
    wa = (((xx - cw) / ((uw - wx) / 2.05)) ** 2 + ((yy - uh / 2 - 3) / (uh / 2.2)) ** 2 <= 1).astype(np.uint8) * 255
    ba[:, bx:] = 0
    wa[:, :wx] = 0
    return black, ba, white, wa


def _bbox(a):
    ys, xs = np.nonzero(a)
    if len(xs) == 0:
        return (0, 0, 0, 0)
    return (int(xs.min()), int(ys.min()), int(xs.max()) + 1, int(ys.max()) + 1)


def _gpu_step(enbl, ctx, black, ba, bbb, white, wa, wbb, itype, exact_levels=0, seed=5):
    uh, uw, C = black.shape
    d_b, d_ba, d_w, d_wa = _dev(black), _dev(ba), _dev(white), _dev(wa)
    bb = enbl.Rect(*bbb)
    wb = enbl.Rect(*wbb)
    outcome = ctypes.c_int(-1)
    torch.cuda.synchronize()
    rc = enbl.lib().enbl_enblend_step(ctx.h, itype, C, uw, uh, 0, exact_levels, 1, 2, _vp(d_b), uw * C, _vp(d_ba), uw,
                                      ctypes.byref(bb), _vp(d_w), uw * C, _vp(d_wa), uw, ctypes.byref(wb), 1, seed,
                                      ctypes.byref(outcome))
    ctx.sync()
    return rc, outcome.value, d_b.cpu().numpy(), d_ba.cpu().numpy(), bb.tuple()


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("geom", [(401, 173, 251, 163), (1200, 700, 800, 403)])
def test_enblend_step_matches_oracle(enbl, gpu_ctx, oracle, itype, geom):
    uw, uh, bx, wx = geom
    print(f"enblend step itype={itype} canvas {uw}x{uh} black<{bx} white>={wx}", flush=True)
    rng = np.random.default_rng(uw + itype)
    black, ba, white, wa = _pair_scene(rng, itype, uw, uh, bx, wx)
    bbb, wbb = _bbox(ba), _bbox(wa)
    erc, eout, eb, eba, ebb, _ = oracle.enblend_step(black, ba, bbb, white, wa, wbb, itype, seed=5)
    assert erc == 0 and eout == 0
    rc, out, b, bal, nbb = _gpu_step(enbl, gpu_ctx, black, ba, bbb, white, wa, wbb, itype)
    enbl.check(rc)
    assert out == eout and nbb == ebb
    np.testing.assert_array_equal(bal, eba)
    np.testing.assert_array_equal(b, eb)


@pytest.mark.parametrize("itype", [U8, U16, F32])
def test_enblend_step_grayscale_matches_oracle(enbl, gpu_ctx, oracle, itype):
    """One-channel layers through the loop step (NFT mask, ROI, blend): the
    mask pyramid must not alias an image pyramid of the same type."""
    uw, uh, bx, wx = 1200, 700, 800, 403
    print(f"enblend step C=1 itype={itype} canvas {uw}x{uh}", flush=True)
    rng = np.random.default_rng(90 + itype)
    black, ba, white, wa = _pair_scene(rng, itype, uw, uh, bx, wx, C=1)
    bbb, wbb = _bbox(ba), _bbox(wa)
    erc, eout, eb, eba, ebb, _ = oracle.enblend_step(black, ba, bbb, white, wa, wbb, itype, seed=5)
    assert erc == 0 and eout == 0
    rc, out, b, bal, nbb = _gpu_step(enbl, gpu_ctx, black, ba, bbb, white, wa, wbb, itype)
    enbl.check(rc)
    assert out == eout and nbb == ebb
    np.testing.assert_array_equal(bal, eba)
    np.testing.assert_array_equal(b, eb)


def test_enblend_step_redundant_disjoint_levels(enbl, gpu_ctx, oracle):
    rng = np.random.default_rng(17)
    black, ba, white, wa = _pair_scene(rng, U16, 300, 120, 180, 120)
    cases = []
    wa2 = np.zeros_like(wa)
    wa2[40:60, 50:90] = 255                      # inside black: redundant
    cases.append((wa2, 0))
    wa3 = np.zeros_like(wa)
    wa3[5:30, 250:290] = 255                     # disjoint: copied
    cases.append((wa3, 0))
    cases.append((wa3, 4))                       # disjoint with --levels=4: blended
    cases.append((wa3, -1))
    for wa_, levels in cases:
        print(f"step case levels={levels} white box={_bbox(wa_)}", flush=True)
        erc, eout, eb, eba, ebb, _ = oracle.enblend_step(black, ba, _bbox(ba), white, wa_, _bbox(wa_), U16,
                                                         exact_levels=levels)
        rc, out, b, bal, nbb = _gpu_step(enbl, gpu_ctx, black, ba, _bbox(ba), white, wa_, _bbox(wa_), U16,
                                         exact_levels=levels)
        assert rc == erc == 0
        assert out == eout and nbb == ebb
        np.testing.assert_array_equal(bal, eba)
        np.testing.assert_array_equal(b, eb)


@pytest.mark.parametrize("itype", [U8, U16, F32])
def test_enblend_step_with_wraparound_matches_oracle(enbl, gpu_ctx, oracle, itype):
    """A union box as wide as the canvas with wraparound: the strip seam mask,
    the ROI across the whole width and the blend with wraparound."""
    rng = np.random.default_rng(60 + itype)
    uw, uh = 640, 211
    black = _samples(rng, uh, uw, 3, itype)
    white = _samples(rng, uh, uw, 3, itype)
    ba = np.zeros((uh, uw), np.uint8)
    wa = np.zeros((uh, uw), np.uint8)
    ba[3:uh - 2, : int(0.62 * uw)] = 255
    wa[1:uh - 4, int(0.40 * uw):] = 255
    print(f"wraparound step itype={itype} canvas {uw}x{uh}", flush=True)
    erc, eout, eb, eba, ebb, _ = oracle.enblend_step(black, ba, _bbox(ba), white, wa, _bbox(wa), itype, wrap=True,
                                                     seed=5)
    assert erc == 0 and eout == 0
    d_b, d_ba, d_w, d_wa = _dev(black), _dev(ba), _dev(white), _dev(wa)
    bb, wb = enbl.Rect(*_bbox(ba)), enbl.Rect(*_bbox(wa))
    outcome = ctypes.c_int(-1)
    torch.cuda.synchronize()
    enbl.check(enbl.lib().enbl_enblend_step(gpu_ctx.h, itype, 3, uw, uh, 1, 0, 1, 2, _vp(d_b), uw * 3, _vp(d_ba), uw,
                                            ctypes.byref(bb), _vp(d_w), uw * 3, _vp(d_wa), uw, ctypes.byref(wb), 1, 5,
                                            ctypes.byref(outcome)))
    gpu_ctx.sync()
    assert outcome.value == 0 and bb.tuple() == ebb
    np.testing.assert_array_equal(d_ba.cpu().numpy(), eba)
    np.testing.assert_array_equal(d_b.cpu().numpy(), eb)


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("pre_assemble", [False, True])
def test_enblend_layers_matches_oracle_loop(enbl, gpu_ctx, oracle, itype, pre_assemble):
    """enblendMain's loop over five overlapping layers, against the oracle's
    assemble + step sequence with seed + m."""
    print(f"enblend layers itype={itype} pre_assemble={pre_assemble}", flush=True)
    rng = np.random.default_rng(300 + itype + pre_assemble)
    uw, uh = 700, 260
    specs = [(240, 260, 0, 10), (240, 250, 200, 0), (200, 240, 420, 40), (100, 120, 580, 150), (90, 100, 10, 160)]
    layers = [(_layer(rng, h, w, itype), x, y, itype) for h, w, x, y in specs]
    seed = 11
    taken = [0] * len(layers)
    _, black, ba, bbb, taken = oracle.assemble(layers, itype, 3, uw, uh, taken, not pre_assemble)
    counts = [0, 0, 0]
    m = 0
    while not all(taken):
        _, white, wa, wbb, taken = oracle.assemble(layers, itype, 3, uw, uh, taken, not pre_assemble)
        rc, out, black, ba, bbb, _ = oracle.enblend_step(black, ba, bbb, white, wa, wbb, itype, seed=seed + m)
        assert rc == 0
        counts[out] += 1
        m += 1
    dl = [(_dev(a), x, y, ft) for a, x, y, ft in layers]
    tab = enbl.layer_table(dl)
    d_b = torch.zeros((uh, uw * 3), dtype=getattr(torch, np.dtype(NP[itype]).name), device="cuda")
    d_ba = torch.zeros((uh, uw), dtype=torch.uint8, device="cuda")
    bb = enbl.Rect()
    cnt = (ctypes.c_int * 3)()
    torch.cuda.synchronize()
    enbl.check(enbl.lib().enbl_enblend_layers(gpu_ctx.h, itype, 3, uw, uh, len(layers), tab, int(pre_assemble), 0, 0,
                                              1, 2, 1, seed, _vp(d_b), uw * 3, _vp(d_ba), uw, ctypes.byref(bb), cnt))
    assert list(cnt) == counts and bb.tuple() == bbb
    np.testing.assert_array_equal(d_ba.cpu().numpy(), ba)
    np.testing.assert_array_equal(d_b.cpu().numpy().reshape(uh, uw, 3), black)


@pytest.mark.parametrize("itype,otype", [(U8, U8), (U16, U16), (F32, F32), (U16, U8), (U8, F32), (U16, F32),
                                         (F32, U8), (F32, U16)])
@pytest.mark.parametrize("C", [3, 1])
def test_checkpoint_matches_oracle(enbl, gpu_ctx, oracle, itype, otype, C):
    rng = np.random.default_rng(itype * 100 + otype + C)
    h, w = 67, 301
    img = _samples(rng, h, w, C, itype)
    if itype == F32:
        img[0, :8, 0] = [np.nan, np.inf, -np.inf, -0.5, 2.0, 1.0, 0.5, 1e-9]
    al = (rng.random((h, w)) < 0.6).astype(np.uint8) * rng.integers(1, 256, size=(h, w)).astype(np.uint8)
    eout, ect = oracle.checkpoint(img, al, itype, otype)
    d_i, d_a = _dev(img), _dev(al)
    opitch = w * (C + 1) + 2
    tdt = {U8: torch.uint8, U16: torch.uint16, F32: torch.float32, U32: torch.uint32}[ect]
    d_o = torch.zeros((h, opitch), dtype=tdt, device="cuda")
    ct = ctypes.c_int(-1)
    torch.cuda.synchronize()
    enbl.check(enbl.lib().enbl_checkpoint(gpu_ctx.h, itype, C, w, h, _vp(d_i), w * C, _vp(d_a), w, otype, _vp(d_o),
                                          opitch, ctypes.byref(ct)))
    gpu_ctx.sync()
    assert ct.value == ect
    out = d_o.cpu().numpy()
    assert (out[:, w * (C + 1):] == 0).all()
    np.testing.assert_array_equal(out[:, :w * (C + 1)].reshape(h, w, C + 1), eout)


def test_checkpoint_rejects_overflowing_conversion(enbl, gpu_ctx):
    d = torch.zeros((2, 6), dtype=torch.uint8, device="cuda")
    o = torch.zeros((2, 8), dtype=torch.uint16, device="cuda")
    ct = ctypes.c_int(-1)
    rc = enbl.lib().enbl_checkpoint(gpu_ctx.h, U8, 3, 2, 2, _vp(d), 6, _vp(d), 2, U16, _vp(o), 8, ctypes.byref(ct))
    assert rc == enbl.ENBL_E_UNSUPPORTED


@pytest.mark.parametrize("itype,otype", [(U8, U8), (U16, U16), (U16, U8), (F32, U16)])
def test_tiff_layers_to_tiff_output(enbl, gpu_ctx, oracle, itype, otype, tmp_path):
    """enblend's whole file path: layers in one multi-directory TIFF (each with
    its canvas position), decoded on the host, blended on the device, the
    checkpoint conversion, written as TIFF and read back -- against the
    oracle's loop and checkpoint over the same decoded layers."""
    print(f"tiff pipeline itype={itype} otype={otype}", flush=True)
    rng = np.random.default_rng(700 + itype + otype)
    uw, uh = 520, 200
    specs = [(200, 260, 0, 0), (190, 250, 200, 8), (120, 150, 360, 60)]
    src = str(tmp_path / "layers.tif")
    enbl.tiff_write(src, [(_layer(rng, h, w, itype), 1, x, y, uw, uh) for h, w, x, y in specs], enbl.TIFF_LZW)
    layers = []
    for k in range(enbl.tiff_info(src)["pages"]):
        a, info = enbl.tiff_read(src, k)
        assert (info["canvas_w"], info["canvas_h"], info["ftype"]) == (uw, uh, itype)
        layers.append((a, info["x"], info["y"], info["ftype"]))
    seed = 3
    taken = [0] * len(layers)
    _, black, ba, bbb, taken = oracle.assemble(layers, itype, 3, uw, uh, taken, False)
    m = 0
    while not all(taken):
        _, white, wa, wbb, taken = oracle.assemble(layers, itype, 3, uw, uh, taken, False)
        rc, _, black, ba, bbb, _ = oracle.enblend_step(black, ba, bbb, white, wa, wbb, itype, seed=seed + m)
        assert rc == 0
        m += 1
    eout, ect = oracle.checkpoint(black, ba, itype, otype)

    dl = [(_dev(a), x, y, ft) for a, x, y, ft in layers]   # kept alive while the device reads them
    tab = enbl.layer_table(dl)
    d_b = torch.zeros((uh, uw * 3), dtype=getattr(torch, np.dtype(NP[itype]).name), device="cuda")
    d_ba = torch.zeros((uh, uw), dtype=torch.uint8, device="cuda")
    bb = enbl.Rect()
    torch.cuda.synchronize()
    enbl.check(enbl.lib().enbl_enblend_layers(gpu_ctx.h, itype, 3, uw, uh, len(layers), tab, 1, 0, 0, 1, 2, 1, seed,
                                              _vp(d_b), uw * 3, _vp(d_ba), uw, ctypes.byref(bb), None))
    tdt = {U8: torch.uint8, U16: torch.uint16, F32: torch.float32, U32: torch.uint32}[ect]
    d_o = torch.zeros((uh, uw * 4), dtype=tdt, device="cuda")
    ct = ctypes.c_int(-1)
    enbl.check(enbl.lib().enbl_checkpoint(gpu_ctx.h, itype, 3, uw, uh, _vp(d_b), uw * 3, _vp(d_ba), uw, otype,
                                          _vp(d_o), uw * 4, ctypes.byref(ct)))
    gpu_ctx.sync()
    assert ct.value == ect
    out = d_o.cpu().numpy().reshape(uh, uw, 4)
    dst = str(tmp_path / "out.tif")
    enbl.tiff_write(dst, [(out, 1, 0, 0)], enbl.TIFF_DEFLATE)
    back, info = enbl.tiff_read(dst)
    assert info["extra"] == 1 and info["ftype"] == ect
    np.testing.assert_array_equal(back, eout)
