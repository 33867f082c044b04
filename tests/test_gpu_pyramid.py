"""GPU parity of the HIP pyramid kernels against the oracle, through the C ABI.

Integer pyramids and every double pyramid level computed from identical
inputs must be bit-exact; float images go through the device log1p/expm1 and
are compared within 2e-6 relative.
"""
import ctypes

import numpy as np
import pytest

from oracle_binding import F32, F64, I16, I32, NP_P, U8, U16

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev_planes(arr, pitch_mult=64):
    """(C, h, w) numpy -> device tensor (C, h, pitch) and its pitch/plane."""
    C, h, w = arr.shape
    pitch = (w + pitch_mult - 1) // pitch_mult * pitch_mult
    t = torch.zeros((C, h, pitch), dtype=getattr(torch, str(arr.dtype)), device="cuda")
    t[:, :, :w] = torch.from_numpy(np.ascontiguousarray(arr)).cuda()
    return t, pitch, h * pitch


def _dev_alpha(a):
    h, w = a.shape
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t, w


def _vp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _gpu_reduce(enbl, ctx, x, alpha, wrap, ptype):
    C, sh, sw = x.shape
    dh, dw = (sh + 1) // 2, (sw + 1) // 2
    src, sp, spl = _dev_planes(x)
    dst, dp, dpl = _dev_planes(np.zeros((C, dh, dw), x.dtype))
    sa = da = None
    sap = dap = 0
    if alpha is not None:
        sa, sap = _dev_alpha(alpha)
        da = torch.zeros((dh, dw), dtype=torch.uint8, device="cuda")
        dap = dw
    torch.cuda.synchronize()
    enbl.check(enbl.lib().enbl_reduce(ctx.h, ptype, C, int(wrap), _vp(src), sw, sh, sp, spl, _vp(sa), sap,
                                      _vp(dst), dw, dh, dp, dpl, _vp(da), dap))
    ctx.sync()
    out = dst[:, :, :dw].cpu().numpy()
    return (out, da.cpu().numpy()) if alpha is not None else out


def _gpu_expand(enbl, ctx, x, d, add, wrap, ptype):
    C, sh, sw = x.shape
    _, dh, dw = d.shape
    src, sp, spl = _dev_planes(x)
    dst, dp, dpl = _dev_planes(d)
    torch.cuda.synchronize()
    enbl.check(enbl.lib().enbl_expand(ctx.h, ptype, C, int(add), int(wrap), _vp(src), sw, sh, sp, spl,
                                      _vp(dst), dw, dh, dp, dpl))
    ctx.sync()
    return dst[:, :, :dw].cpu().numpy()


def _rand(rng, ptype, shape, signed=False):
    if ptype == I16:
        lo = -16000 if signed else 0
        return rng.integers(lo, 32641, size=shape).astype(np.int16)
    if ptype == I32:
        lo = -(1 << 23) if signed else 0
        return rng.integers(lo, 65535 * 128 + 1, size=shape).astype(np.int32)
    v = rng.random(shape) * 4.0 + 1.0
    return (-v if signed else v).astype(np.float64)


REDUCE_SHAPES = [(2, 2), (2, 3), (3, 2), (3, 3), (5, 4), (4, 5), (7, 9), (16, 16), (17, 33), (33, 17),
                 (64, 130), (125, 127), (126, 250), (131, 124), (61, 300), (300, 61), (513, 257)]


@pytest.fixture(params=[1, 2], ids=["direct", "streaming"])
def policy(request, enbl):
    enbl.check(enbl.lib().enbl_set_kernel_policy(request.param))
    yield request.param
    enbl.check(enbl.lib().enbl_set_kernel_policy(0))


@pytest.mark.parametrize("ptype", [I16, I32, F64])
@pytest.mark.parametrize("C", [1, 3])
@pytest.mark.parametrize("wrap", [False, True])
@pytest.mark.parametrize("use_alpha", [True, False])
def test_reduce_matches_oracle(enbl, gpu_ctx, oracle, ptype, C, wrap, use_alpha, policy):
    rng = np.random.default_rng(100 + ptype * 8 + C * 2 + wrap + 16 * use_alpha)
    for (h, w) in REDUCE_SHAPES:
        print(f"reduce ptype={ptype} C={C} wrap={wrap} alpha={use_alpha} {w}x{h}", flush=True)
        x = _rand(rng, ptype, (C, h, w))
        a = np.where(rng.random((h, w)) < 0.75, 255, 0).astype(np.uint8) if use_alpha else None
        got = _gpu_reduce(enbl, gpu_ctx, x, a, wrap, ptype)
        for c in range(C):
            want = oracle.reduce(x[c], a, wrap, ptype)
            if use_alpha:
                np.testing.assert_array_equal(got[0][c], want[0], err_msg=f"{w}x{h} c{c}")
                np.testing.assert_array_equal(got[1], want[1])
            else:
                np.testing.assert_array_equal(got[c], want, err_msg=f"{w}x{h} c{c}")


EXPAND_SHAPES = [(1, 1, 1, 2), (1, 1, 2, 1), (1, 1, 2, 2), (1, 3, 2, 5), (1, 3, 1, 6), (3, 1, 6, 2), (3, 1, 5, 1),
                 (2, 2, 3, 3), (2, 2, 4, 4), (3, 5, 5, 10), (3, 5, 6, 9), (9, 13, 17, 26), (33, 65, 66, 129),
                 (63, 62, 125, 124), (64, 64, 128, 127), (100, 130, 199, 259), (257, 129, 514, 257)]


@pytest.mark.parametrize("ptype", [I16, I32, F64])
@pytest.mark.parametrize("C", [1, 3])
@pytest.mark.parametrize("wrap", [False, True])
@pytest.mark.parametrize("add", [False, True])
def test_expand_matches_oracle(enbl, gpu_ctx, oracle, ptype, C, wrap, add, policy):
    rng = np.random.default_rng(200 + ptype * 8 + C * 2 + wrap + 16 * add)
    for (sh, sw, dh, dw) in EXPAND_SHAPES:
        print(f"expand ptype={ptype} C={C} wrap={wrap} add={add} src {sw}x{sh} dst {dw}x{dh}", flush=True)
        x = _rand(rng, ptype, (C, sh, sw), signed=True)
        d = _rand(rng, ptype, (C, dh, dw), signed=True)
        got = _gpu_expand(enbl, gpu_ctx, x, d, add, wrap, ptype)
        for c in range(C):
            want = oracle.expand(x[c], d[c], add, wrap, ptype)
            np.testing.assert_array_equal(got[c], want, err_msg=f"src {sw}x{sh} dst {dw}x{dh} c{c}")


def test_reduce_precondition_error(enbl, gpu_ctx):
    x = torch.zeros((1, 4, 64), dtype=torch.int32, device="cuda")
    y = torch.zeros((1, 2, 64), dtype=torch.int32, device="cuda")
    rc = enbl.lib().enbl_reduce(gpu_ctx.h, I32, 1, 0, _vp(x), 1, 4, 64, 256, None, 0, _vp(y), 1, 2, 64, 128, None, 0)
    assert rc == enbl.ENBL_E_PRECONDITION
    rc = enbl.lib().enbl_reduce(gpu_ctx.h, I32, 1, 0, _vp(x), 8, 4, 64, 256, None, 0, _vp(y), 3, 2, 64, 128, None, 0)
    assert rc == enbl.ENBL_E_SIZE


def _img(rng, itype, h, w, C=3):
    if itype == U8:
        return rng.integers(0, 256, size=(h, w, C)).astype(np.uint8)
    if itype == U16:
        return rng.integers(0, 65536, size=(h, w, C)).astype(np.uint16)
    return (np.exp(rng.normal(0, 2, size=(h, w, C)))).astype(np.float32)


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("wrap", [False, True])
def test_pyramids_match_oracle(enbl, gpu_ctx, oracle, itype, wrap, policy):
    rng = np.random.default_rng(300 + itype + wrap)
    ptype = {U8: I16, U16: I32, F32: F64}[itype]
    frac = 0 if itype == F32 else 7
    for (h, w) in [(67, 91), (128, 128), (250, 333)]:
        L = oracle.auto_levels(w, h)
        print(f"pyramid itype={itype} wrap={wrap} {w}x{h} L={L}", flush=True)
        img = _img(rng, itype, h, w)
        a = np.where(rng.random((h, w)) < 0.8, 255, 0).astype(np.uint8)
        dimg = torch.from_numpy(img).cuda()
        da = torch.from_numpy(a).cuda()
        for lap in (False, True):
            p = enbl.Pyramid(gpu_ctx, ptype, 3, w, h, L)
            fn = enbl.lib().enbl_laplacian_pyramid if lap else enbl.lib().enbl_gaussian_pyramid
            torch.cuda.synchronize()
            enbl.check(fn(gpu_ctx.h, int(wrap), itype, frac, _vp(dimg), w * 3, _vp(da), w, p.h))
            gpu_ctx.sync()
            want = oracle.pyramid(img, a, L, wrap, lap, itype)
            if itype == F32:
                # level 0 is 1 + log1p(v) (or 1 / (1 - v)) from the device's
                # log1p, which may differ from glibc's in the last bit; every
                # later level follows the double reduce / expand in the same
                # operation order (bit-exact from identical inputs:
                # test_double_pyramid_ops_bit_exact), so the levels agree to
                # a few ulps of the level-0 magnitude.
                lv0 = oracle.to_pyramid(img, F32, F64, 0)
                scale = float(np.abs(lv0).max())
                for l in range(L):
                    got = p.read(l)
                    if not lap:
                        np.testing.assert_allclose(got, want[l], rtol=2e-15 * (l + 1), atol=0,
                                                   err_msg=f"gaussian level {l}")
                    else:
                        np.testing.assert_allclose(got, want[l], rtol=0, atol=4e-15 * scale * (l + 1),
                                                   err_msg=f"laplacian level {l}")
                if lap:
                    enbl.check(enbl.lib().enbl_collapse_pyramid(gpu_ctx.h, int(wrap), p.h))
                    gpu_ctx.sync()
                    np.testing.assert_allclose(p.read(0), lv0, rtol=0, atol=4e-15 * scale * L)
                p.close()
                continue
            for l in range(L):
                np.testing.assert_array_equal(p.read(l), want[l], err_msg=f"level {l} lap={lap}")
            if lap:
                enbl.check(enbl.lib().enbl_collapse_pyramid(gpu_ctx.h, int(wrap), p.h))
                gpu_ctx.sync()
                np.testing.assert_array_equal(p.read(0), img.transpose(2, 0, 1).astype(np.int64) << 7)
            p.close()


@pytest.mark.parametrize("wrap", [False, True])
def test_double_pyramid_ops_bit_exact(enbl, gpu_ctx, oracle, wrap, policy):
    """Reduce / expand chains on a double pyramid with an identical level 0
    are bit-identical to the oracle (same operation order, no contraction)."""
    rng = np.random.default_rng(400 + wrap)
    h, w, L = 150, 203, 5
    lv0 = 1.0 + np.log1p(np.exp(rng.normal(0, 2, size=(3, h, w))))
    a = np.where(rng.random((h, w)) < 0.8, 255, 0).astype(np.uint8)
    levels = [lv0]
    alphas = [a]
    for l in range(1, L):
        nxt = []
        na = None
        for c in range(3):
            r, na = oracle.reduce(levels[-1][c], alphas[-1], wrap, F64)
            nxt.append(r)
        levels.append(np.stack(nxt))
        alphas.append(na)
    for l in range(1, L):
        got = _gpu_reduce(enbl, gpu_ctx, levels[l - 1], alphas[l - 1], wrap, F64)
        np.testing.assert_array_equal(got[0], levels[l], err_msg=f"reduce level {l}")
        d = levels[l - 1]
        got = _gpu_expand(enbl, gpu_ctx, levels[l], d, False, wrap, F64)
        want = np.stack([oracle.expand(levels[l][c], d[c], False, wrap, F64) for c in range(3)])
        np.testing.assert_array_equal(got, want, err_msg=f"expand level {l}")


@pytest.mark.parametrize("itype", [U8, U16])
def test_blend_level_matches_oracle(enbl, gpu_ctx, oracle, itype):
    rng = np.random.default_rng(500 + itype)
    ptype = I16 if itype == U8 else I32
    mtype = I16 if itype == U8 else I32
    mfrac = 7 if itype == U8 else 15
    h, w = 97, 131
    L = 3
    mask = rng.integers(0, 256, size=(h, w)).astype(np.uint8)
    mask[rng.random((h, w)) < 0.3] = 0
    mask[rng.random((h, w)) < 0.3] = 255
    wimg, bimg = _img(rng, itype, h, w), _img(rng, itype, h, w)
    full = np.full((h, w), 255, np.uint8)
    mgp = enbl.Pyramid(gpu_ctx, mtype, 1, w, h, L)
    wlp = enbl.Pyramid(gpu_ctx, ptype, 3, w, h, L)
    blp = enbl.Pyramid(gpu_ctx, ptype, 3, w, h, L)
    dm, dw_, db, dfull = (torch.from_numpy(v).cuda() for v in (mask, wimg, bimg, full))
    torch.cuda.synchronize()
    enbl.check(enbl.lib().enbl_gaussian_pyramid(gpu_ctx.h, 0, U8, mfrac, _vp(dm), w, None, 0, mgp.h))
    enbl.check(enbl.lib().enbl_laplacian_pyramid(gpu_ctx.h, 0, itype, 7, _vp(dw_), 3 * w, _vp(dfull), w, wlp.h))
    enbl.check(enbl.lib().enbl_laplacian_pyramid(gpu_ctx.h, 0, itype, 7, _vp(db), 3 * w, _vp(dfull), w, blp.h))
    gpu_ctx.sync()
    m_levels = [mgp.read(l)[0] for l in range(L)]
    w_levels = [wlp.read(l) for l in range(L)]
    b_levels = [blp.read(l) for l in range(L)]
    wv = float(255 << mfrac)
    enbl.check(enbl.lib().enbl_blend(gpu_ctx.h, mgp.h, wlp.h, blp.h, wv))
    gpu_ctx.sync()
    for l in range(L):
        want = oracle.blend_level(m_levels[l], w_levels[l], b_levels[l], mtype, ptype, wv)
        np.testing.assert_array_equal(blp.read(l), want, err_msg=f"level {l}")


def _enblend_inputs(rng, itype, h, w, C=3):
    white, black = _img(rng, itype, h, w, C), _img(rng, itype, h, w, C)
    yy, xx = np.mgrid[0:h, 0:w]
    wa = np.where((xx - 0.3 * w) ** 2 / (0.45 * w) ** 2 + (yy - 0.5 * h) ** 2 / (0.48 * h) ** 2 < 1, 255, 0).astype(np.uint8)
    ba = np.where((xx - 0.7 * w) ** 2 / (0.45 * w) ** 2 + (yy - 0.5 * h) ** 2 / (0.48 * h) ** 2 < 1, 255, 0).astype(np.uint8)
    mask = np.where(xx < 0.5 * w + 0.1 * (yy - h / 2), 0, 255).astype(np.uint8)
    return white, wa, black, ba, mask


@pytest.mark.parametrize("itype", [U8, U16, F32])
def test_enblend_pair_small_level_launches(enbl, gpu_ctx, oracle, itype):
    """Automatic kernel choice on a size whose level 1 takes the fused pass
    (over 2^20 pixels) and whose smaller levels run in the shared multi-job
    launches (white, black and mask reduces with expand-subtracts, blend,
    blend + collapse); odd sizes, compared with the oracle."""
    rng = np.random.default_rng(640 + itype)
    h, w = 1501, 2899
    L = oracle.auto_levels(w, h)
    print(f"enblend_pair auto itype={itype} {w}x{h} L={L}", flush=True)
    white, wa, black, ba, mask = _enblend_inputs(rng, itype, h, w)
    want, want_a = oracle.enblend_pair(white, wa, black, ba, mask, L, False, itype, 1, 77)
    dwh, dwa, dbk, dba, dm = (torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in (white, wa, black, ba, mask))
    torch.cuda.synchronize()
    enbl.check(enbl.lib().enbl_enblend_pair(gpu_ctx.h, itype, 3, w, h, L, 0,
                                            _vp(dwh), 3 * w, _vp(dwa), w, _vp(dbk), 3 * w, _vp(dba), w,
                                            _vp(dm), w, 1, 77))
    gpu_ctx.sync()
    got, got_a = dbk.cpu().numpy(), dba.cpu().numpy()
    np.testing.assert_array_equal(got_a, want_a)
    if itype == F32:
        sel = want_a != 0
        np.testing.assert_allclose(got[sel], want[sel], rtol=2e-6, atol=0)
    else:
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("wrap", [False, True])
def test_enblend_pair_matches_oracle(enbl, gpu_ctx, oracle, itype, wrap, policy):
    rng = np.random.default_rng(600 + itype + wrap)
    for (h, w) in [(96, 160), (301, 457)]:
        L = oracle.auto_levels(w, h)
        print(f"enblend_pair itype={itype} wrap={wrap} {w}x{h} L={L}", flush=True)
        white, wa, black, ba, mask = _enblend_inputs(rng, itype, h, w)
        want, want_a = oracle.enblend_pair(white, wa, black, ba, mask, L, wrap, itype, 1, 77)
        dwh, dwa, dbk, dba, dm = (torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in (white, wa, black, ba, mask))
        torch.cuda.synchronize()
        enbl.check(enbl.lib().enbl_enblend_pair(gpu_ctx.h, itype, 3, w, h, L, int(wrap),
                                                _vp(dwh), 3 * w, _vp(dwa), w, _vp(dbk), 3 * w, _vp(dba), w,
                                                _vp(dm), w, 1, 77))
        gpu_ctx.sync()
        got, got_a = dbk.cpu().numpy(), dba.cpu().numpy()
        np.testing.assert_array_equal(got_a, want_a)
        if itype == F32:
            sel = want_a != 0
            np.testing.assert_allclose(got[sel], want[sel], rtol=2e-6, atol=0)
        else:
            np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("shape", [(96, 160), (301, 457), (1501, 2899)])
def test_enblend_pair_grayscale_matches_oracle(enbl, gpu_ctx, oracle, itype, shape):
    """One-channel images: the mask Gaussian pyramid and the white Laplacian
    pyramid share type and channel count here (Int16 for 8-bit, Int32 for
    16-bit images), so they must still be distinct pyramids.  Sizes cover the
    plain level-by-level path, the shared small-level launches and the fused
    streaming passes; two calls in a row on the same context."""
    h, w = shape
    rng = np.random.default_rng(700 + itype + h)
    L = oracle.auto_levels(w, h)
    for rep in range(2):
        print(f"enblend_pair C=1 itype={itype} {w}x{h} L={L} call {rep}", flush=True)
        white, wa, black, ba, mask = _enblend_inputs(rng, itype, h, w, C=1)
        want, want_a = oracle.enblend_pair(white, wa, black, ba, mask, L, False, itype, 1, 31 + rep)
        dwh, dwa, dbk, dba, dm = (torch.from_numpy(np.ascontiguousarray(v)).cuda()
                                  for v in (white, wa, black, ba, mask))
        torch.cuda.synchronize()
        enbl.check(enbl.lib().enbl_enblend_pair(gpu_ctx.h, itype, 1, w, h, L, 0,
                                                _vp(dwh), w, _vp(dwa), w, _vp(dbk), w, _vp(dba), w,
                                                _vp(dm), w, 1, 31 + rep))
        gpu_ctx.sync()
        got, got_a = dbk.cpu().numpy(), dba.cpu().numpy()
        np.testing.assert_array_equal(got_a, want_a)
        if itype == F32:
            sel = want_a != 0
            np.testing.assert_allclose(got[sel], want[sel], rtol=2e-6, atol=0)
        else:
            np.testing.assert_array_equal(got, want)
