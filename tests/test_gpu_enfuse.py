"""GPU parity of enfuse weights and fusion against the oracle (C ABI)."""
import ctypes

import numpy as np
import pytest

from oracle_binding import F32, U8, U16, enfuse_params as orc_params

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _stack(rng, itype, K, h, w, scene=True):
    """K exposures of one smooth scene (so weights vary smoothly)."""
    yy, xx = np.mgrid[0:h, 0:w]
    base = 0.5 + 0.4 * np.sin(xx / 7.0) * np.cos(yy / 5.0)
    tint = rng.uniform(0.7, 1.3, size=3)
    imgs = []
    for k in range(K):
        ev = (k - (K - 1) / 2) * 1.0
        rad = base[..., None] * tint[None, None, :] * 2.0 ** ev
        if itype == F32:
            imgs.append((rad + rng.normal(0, 1e-3, rad.shape)).astype(np.float32))
        else:
            m = 255 if itype == U8 else 65535
            v = np.clip(rad, 0, 1) ** (1 / 2.2) * m + rng.normal(0, 1, rad.shape)
            imgs.append(np.clip(np.rint(v), 0, m).astype(np.uint8 if itype == U8 else np.uint16))
    return imgs


@pytest.fixture(params=[0, 2], ids=["auto", "streaming"])
def policy(request, enbl):
    enbl.check(enbl.lib().enbl_set_kernel_policy(request.param))
    yield request.param
    enbl.check(enbl.lib().enbl_set_kernel_policy(0))


CASES = [
    dict(w_exposure=1.0, w_saturation=0.2),
    dict(w_exposure=1.0, w_saturation=0.2, w_contrast=0.5, w_entropy=0.3, contrast_window=5, entropy_window=3),
    dict(w_exposure=0.5, w_contrast=1.0, w_entropy=0.3, contrast_window=3, entropy_window=7, exposure_function=2),
]


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_enfuse_mask_matches_oracle(enbl, gpu_ctx, oracle, itype, case):
    rng = np.random.default_rng(700 + itype + case)
    h, w = 61, 83
    img = _stack(rng, itype, 1, h, w)[0]
    a = np.where(rng.random((h, w)) < 0.9, 255, 0).astype(np.uint8)
    kw = CASES[case]
    want = oracle.enfuse_mask(img, a, itype, orc_params(**kw))
    dimg, da = torch.from_numpy(img).cuda(), torch.from_numpy(a).cuda()
    dm = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    print(f"enfuse_mask itype={itype} case={case}", flush=True)
    enbl.check(enbl.lib().enbl_enfuse_mask(gpu_ctx.h, itype, w, h, _vp(dimg), 3 * w, _vp(da), w,
                                           ctypes.byref(enbl.enfuse_params(**kw)), _vp(dm), w))
    gpu_ctx.sync()
    got = dm.cpu().numpy()
    if itype == F32:
        np.testing.assert_allclose(got, want, rtol=1e-6, atol=0)
    else:
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_enfuse_matches_oracle(enbl, gpu_ctx, oracle, itype, case, policy):
    rng = np.random.default_rng(800 + itype + case)
    for (h, w, K) in [(64, 96, 3), (133, 101, 4)]:
        imgs = _stack(rng, itype, K, h, w)
        alphas = [np.where(rng.random((h, w)) < 0.95, 255, 0).astype(np.uint8) for _ in range(K)]
        L = oracle.auto_levels(w, h)
        kw = CASES[case]
        print(f"enfuse itype={itype} case={case} {w}x{h} K={K} L={L}", flush=True)
        want, want_a = oracle.enfuse(imgs, alphas, itype, orc_params(**kw), L, dither=1, seed=3)
        dimgs = [torch.from_numpy(i).cuda() for i in imgs]
        das = [torch.from_numpy(a).cuda() for a in alphas]
        out = torch.zeros_like(dimgs[0])
        oa = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        enbl.enfuse(gpu_ctx, itype, dimgs, das, enbl.enfuse_params(**kw), L, out, oa, dither=1, seed=3)
        gpu_ctx.sync()
        np.testing.assert_array_equal(oa.cpu().numpy(), want_a)
        got = out.cpu().numpy()
        if itype == F32:
            sel = want_a != 0
            np.testing.assert_allclose(got[sel], want[sel], rtol=2e-6, atol=0)
        else:
            np.testing.assert_array_equal(got, want)


def _tied_stack(rng, itype, h, w):
    """One image with many repeated sample values (so entropy windows hold
    bins of several samples); float images also mix -0.0 and +0.0."""
    yy, xx = np.mgrid[0:h, 0:w]
    base = 0.5 + 0.45 * np.sin(xx / 9.0) * np.cos(yy / 6.0)
    img = base[..., None] * rng.uniform(0.7, 1.3, size=3)[None, None, :]
    if itype == F32:
        v = np.round(img * 16.0) / 16.0 - 0.25
        v[rng.random(v.shape) < 0.05] = -0.0
        v[rng.random(v.shape) < 0.05] = 0.0
        return v.astype(np.float32)
    m = 255 if itype == U8 else 65535
    v = np.round(img * 24.0) / 24.0 * m
    return np.clip(v, 0, m).astype(np.uint8 if itype == U8 else np.uint16)


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("windows", [(3, 3), (5, 7), (7, 5)])
@pytest.mark.parametrize("holes", [False, True], ids=["full-alpha", "holes"])
def test_contrast_entropy_tiles_bit_exact(enbl, gpu_ctx, oracle, itype, windows, holes):
    """Contrast and entropy weights over several LDS tiles and column chunks
    (150 x 203, window rows past one wave), with full windows (the sorting
    network path) and windows cut by alpha (the insertion-sort path).  With
    the exposure weight off every term is computed identically, so the float
    weights are compared bit for bit."""
    rng = np.random.default_rng(900 + itype + 10 * windows[0] + windows[1] + (1000 if holes else 0))
    h, w = 150, 203
    img = _tied_stack(rng, itype, h, w)
    a = np.full((h, w), 255, np.uint8)
    if holes:
        a[rng.random((h, w)) < 0.15] = 0
        a[40:60, 100:180] = 0
    kw = dict(w_exposure=0.0, w_saturation=0.2, w_contrast=0.5, w_entropy=0.3,
              contrast_window=windows[0], entropy_window=windows[1])
    want = oracle.enfuse_mask(img, a, itype, orc_params(**kw))
    dimg, da = torch.from_numpy(img).cuda(), torch.from_numpy(a).cuda()
    dm = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    print(f"contrast/entropy tiles itype={itype} windows={windows} holes={holes}", flush=True)
    enbl.check(enbl.lib().enbl_enfuse_mask(gpu_ctx.h, itype, w, h, _vp(dimg), 3 * w, _vp(da), w,
                                           ctypes.byref(enbl.enfuse_params(**kw)), _vp(dm), w))
    gpu_ctx.sync()
    np.testing.assert_array_equal(dm.cpu().numpy().view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("K", [3, 7])
def test_hard_mask_enfuse_matches_oracle(enbl, gpu_ctx, oracle, itype, K):
    """--hard-mask fusion (src/enfuse.h:1402-1443) against the oracle, with
    alpha holes, and with all weights zero (the float 255 / K share)."""
    rng = np.random.default_rng(950 + itype + K)
    h, w = 71, 93
    imgs = _stack(rng, itype, K, h, w)
    alphas = [np.where(rng.random((h, w)) < 0.9, 255, 0).astype(np.uint8) for _ in range(K)]
    L = oracle.auto_levels(w, h)
    for kw in (dict(w_exposure=1.0, w_saturation=0.2, hard_mask=True), dict(w_exposure=0.0, hard_mask=True)):
        print(f"hard mask itype={itype} K={K} {kw}", flush=True)
        want, want_a = oracle.enfuse(imgs, alphas, itype, orc_params(**kw), L, dither=1, seed=5)
        dimgs = [torch.from_numpy(i).cuda() for i in imgs]
        das = [torch.from_numpy(a).cuda() for a in alphas]
        out = torch.zeros_like(dimgs[0])
        oa = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        enbl.enfuse(gpu_ctx, itype, dimgs, das, enbl.enfuse_params(**kw), L, out, oa, dither=1, seed=5)
        gpu_ctx.sync()
        np.testing.assert_array_equal(oa.cpu().numpy(), want_a)
        got = out.cpu().numpy()
        if itype == F32:
            sel = want_a != 0
            np.testing.assert_allclose(got[sel], want[sel], rtol=2e-6, atol=0)
        else:
            np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("ewin", [3, 5, 7])
def test_entropy_window_kinds_bit_exact(enbl, gpu_ctx, oracle, itype, ewin):
    """Full windows of one value (one bin), of distinct values (as many bins
    as samples) and with ties, side by side under full alpha; the entropy
    weight alone, compared bit for bit."""
    rng = np.random.default_rng(1100 + itype + ewin)
    h, w = 80, 210
    m = {U8: 255, U16: 65535, F32: 1.0}[itype]
    img = np.empty((h, w, 3), np.float64)
    img[:, :70] = 0.37 * m                                                       # flat
    img[:, 70:140] = rng.permutation(np.linspace(0.05, 0.95, h * 70 * 3)).reshape(h, 70, 3) * m  # distinct
    img[:, 140:] = np.round(rng.random((h, 70, 3)) * 6) / 6 * m                  # ties
    img = img.astype(np.float32) if itype == F32 else np.rint(img).astype(np.uint8 if itype == U8 else np.uint16)
    a = np.full((h, w), 255, np.uint8)
    kw = dict(w_exposure=0.0, w_entropy=1.0, entropy_window=ewin)
    want = oracle.enfuse_mask(img, a, itype, orc_params(**kw))
    dimg, da = torch.from_numpy(img).cuda(), torch.from_numpy(a).cuda()
    dm = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    print(f"entropy window kinds itype={itype} n={ewin}", flush=True)
    enbl.check(enbl.lib().enbl_enfuse_mask(gpu_ctx.h, itype, w, h, _vp(dimg), 3 * w, _vp(da), w,
                                           ctypes.byref(enbl.enfuse_params(**kw)), _vp(dm), w))
    gpu_ctx.sync()
    np.testing.assert_array_equal(dm.cpu().numpy().view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("ewin", [3, 5, 7])
def test_entropy_constant_column_windows_bit_exact(enbl, gpu_ctx, oracle, itype, ewin):
    """Windows whose columns are each constant (values depending on x only):
    one bin per distinct column value of n samples each, including columns
    of equal value (bins of 2n, 3n samples), beside windows where one column
    or one row breaks the pattern, a single-channel flat region (the pixel's
    minimum is 0) and -0 / +0 columns for float images; the entropy weight
    alone, bit for bit against the oracle."""
    rng = np.random.default_rng(1300 + itype + ewin)
    h, w = 96, 260
    m = {U8: 255, U16: 65535, F32: 1.0}[itype]
    xs = np.arange(w)
    colv = np.empty((w, 3))
    colv[:, 0] = (np.floor(xs / 2) % 9) / 9.0           # pairs of equal columns
    colv[:, 1] = rng.permutation(np.linspace(0.1, 0.9, w))  # distinct columns
    colv[:, 2] = (np.floor(xs / 3) % 5) / 5.0           # triples of equal columns
    img = np.repeat(colv[None, :, :], h, axis=0) * m
    img[40:44, 100:160] = rng.random((4, 60, 3)) * m     # rows that break the columns
    img[70, 200:] = 0.5 * m
    img[:, 220:230, 1] = 0.25 * m                          # flat in one channel only
    if itype == F32:
        img = img.astype(np.float32)
        img[:, 150:152, 0] = -0.0                          # -0 and +0 share a bin
        img[:, 152:154, 0] = 0.0
    else:
        img = np.rint(img).astype(np.uint8 if itype == U8 else np.uint16)
    a = np.full((h, w), 255, np.uint8)
    kw = dict(w_exposure=0.0, w_entropy=1.0, entropy_window=ewin)
    want = oracle.enfuse_mask(img, a, itype, orc_params(**kw))
    dimg, da = torch.from_numpy(np.ascontiguousarray(img)).cuda(), torch.from_numpy(a).cuda()
    dm = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    print(f"entropy constant-column windows itype={itype} n={ewin}", flush=True)
    enbl.check(enbl.lib().enbl_enfuse_mask(gpu_ctx.h, itype, w, h, _vp(dimg), 3 * w, _vp(da), w,
                                           ctypes.byref(enbl.enfuse_params(**kw)), _vp(dm), w))
    gpu_ctx.sync()
    np.testing.assert_array_equal(dm.cpu().numpy().view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("ewin", [3, 7])
@pytest.mark.parametrize("fused", [False, True])
def test_entropy_flat_tiles_bit_exact(enbl, gpu_ctx, oracle, itype, ewin, fused):
    """Tiles whose whole window halo has one value in a channel (entropy 0 at
    every pixel without sorting): all channels flat, one channel flat beside
    random ones, -0 / +0 mixed in a flat float region, a flat region broken
    by a single sample or by an alpha hole, and flat regions that end inside
    a tile; the entropy weight alone (fused: with exposure and saturation in
    the same pass), bit for bit against the oracle."""
    rng = np.random.default_rng(1500 + itype + ewin + (7 if fused else 0))
    h, w = 72, 520
    m = {U8: 255, U16: 65535, F32: 1.0}[itype]
    img = rng.random((h, w, 3)) * m
    img[:, 30:200] = 0.4 * m                           # all channels flat over whole tiles
    img[:, 230:330, 1] = 0.7 * m                       # one channel flat
    img[20:60, 350:470] = 0.2 * m
    img[37, 410] = 0.3 * m                             # one sample breaks a flat tile
    img[5:13, 100] = 0.6 * m                           # a short flat-breaking column
    a = np.full((h, w), 255, np.uint8)
    a[50, 150] = 0                                     # an alpha hole inside a flat region
    if itype == F32:
        img = img.astype(np.float32)
        img[:, 60:130, 2] = 0.0
        img[::2, 60:130, 2] = -0.0                     # -0 and +0 share a bin
    else:
        img = np.rint(img).astype(np.uint8 if itype == U8 else np.uint16)
    kw = dict(w_exposure=0.5 if fused else 0.0, w_saturation=0.3 if fused else 0.0, w_entropy=1.0,
              entropy_window=ewin)
    want = oracle.enfuse_mask(img, a, itype, orc_params(**kw))
    dimg, da = torch.from_numpy(np.ascontiguousarray(img)).cuda(), torch.from_numpy(a).cuda()
    dm = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    print(f"entropy flat tiles itype={itype} n={ewin} fused={fused}", flush=True)
    enbl.check(enbl.lib().enbl_enfuse_mask(gpu_ctx.h, itype, w, h, _vp(dimg), 3 * w, _vp(da), w,
                                           ctypes.byref(enbl.enfuse_params(**kw)), _vp(dm), w))
    gpu_ctx.sync()
    np.testing.assert_array_equal(dm.cpu().numpy().view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("ewin", [3, 5, 7])
def test_entropy_ordered_windows_bit_exact(enbl, gpu_ctx, oracle, itype, ewin):
    """Windows of linear ramps: columns strictly monotone with their ranges
    strictly ordered along the row (all samples distinct, no sort), ranges
    that touch (a sample shared by neighbouring columns: tied bins), ramps
    falling to the left or upwards, ramps along y (rows ordered, sorted),
    and ramps broken by one sample or by a plateau; the entropy weight alone,
    bit for bit against the oracle."""
    h, w = 48, 448
    n = ewin
    m = {U8: 255, U16: 65535, F32: 1.0}[itype]
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    step = 1.0 if itype != F32 else 1.0 / 4096
    img = np.empty((h, w, 3), np.float64)
    lo = 0.0 if itype == U8 else 100.0 * step
    img[:, :, 0] = lo + (n * xx + yy) * step * (0.25 if itype == U8 else 1.0)           # ordered columns
    img[:, :, 1] = lo + ((n - 1) * xx + yy) * step * (0.25 if itype == U8 else 1.0)     # touching ranges
    img[:, :, 2] = lo + (n * (w - xx) + (h - yy)) * step * (0.25 if itype == U8 else 1.0)   # falling ramps
    img[:, 200:260, 0] = lo + (xx[:, 200:260] + n * yy[:, 200:260]) * step               # rows ordered
    img[20, 300, :] = img[20, 301, :]                                                   # one sample breaks
    img[:, 330:336, 1] = img[:, 330:331, 1]                                             # a plateau
    if itype == U8:
        img = np.clip(img, 0, 255)
    img = img.astype(np.float32) if itype == F32 else np.rint(img).astype(np.uint8 if itype == U8 else np.uint16)
    a = np.full((h, w), 255, np.uint8)
    kw = dict(w_exposure=0.0, w_entropy=1.0, entropy_window=ewin)
    want = oracle.enfuse_mask(img, a, itype, orc_params(**kw))
    dimg, da = torch.from_numpy(np.ascontiguousarray(img)).cuda(), torch.from_numpy(a).cuda()
    dm = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    print(f"entropy ordered windows itype={itype} n={ewin}", flush=True)
    enbl.check(enbl.lib().enbl_enfuse_mask(gpu_ctx.h, itype, w, h, _vp(dimg), 3 * w, _vp(da), w,
                                           ctypes.byref(enbl.enfuse_params(**kw)), _vp(dm), w))
    gpu_ctx.sync()
    np.testing.assert_array_equal(dm.cpu().numpy().view(np.uint32), want.view(np.uint32))


def _table_call(enbl, ctx, itype, table, entry):
    """Calls one enfuse entry point on a tiny stack with an exposure table;
    returns the status code."""
    h, w, K = 8, 8, 2
    dt = {U8: torch.uint8, U16: torch.int16, F32: torch.float32}[itype]
    imgs = [torch.zeros((h, w, 3), dtype=dt, device="cuda") for _ in range(K)]
    alphas = [torch.full((h, w), 255, dtype=torch.uint8, device="cuda") for _ in range(K)]
    out = torch.zeros((h, w, 3), dtype=dt, device="cuda")
    oa = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
    m = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    prm = enbl.enfuse_params(w_exposure=1.0, w_saturation=0.2, exposure_table=table)
    torch.cuda.synchronize()
    try:
        if entry == "mask":
            enbl.check(enbl.lib().enbl_enfuse_mask(ctx.h, itype, w, h, _vp(imgs[0]), 3 * w, _vp(alphas[0]), w,
                                                   ctypes.byref(prm), _vp(m), w))
        elif entry == "enfuse":
            enbl.enfuse(ctx, itype, imgs, alphas, prm, 2, out, oa)
        else:
            enbl.enfuse_batch(ctx, itype, [(imgs, alphas, out, oa, 2, 0)], prm)
        ctx.sync()
        return enbl.ENBL_OK
    except enbl.EnblError as e:
        return e.code


@pytest.mark.parametrize("entry", ["mask", "enfuse", "batch"])
@pytest.mark.parametrize("itype", [U8, U16])
def test_exposure_table_checked_like_check_weight_function(enbl, gpu_ctx, itype, entry):
    """A user weight table is checked as check_weight_function checks the
    plugin (src/exposure_weight.cc:237-264): a negative weight, a weight above
    one or all weights zero are rejected; zeros elsewhere, denormals counted
    as zero, and NaN (neither negative, above one nor zero) pass."""
    n = 256 if itype == U8 else 65536
    ok = np.linspace(0.0, 1.0, n)
    print(f"exposure table checks itype={itype} entry={entry}", flush=True)
    assert _table_call(enbl, gpu_ctx, itype, ok, entry) == enbl.ENBL_OK
    bad = ok.copy()
    bad[n // 3] = -1e-300
    assert _table_call(enbl, gpu_ctx, itype, bad, entry) == enbl.ENBL_E_PRECONDITION
    bad = ok.copy()
    bad[-1] = np.nextafter(1.0, 2.0)
    assert _table_call(enbl, gpu_ctx, itype, bad, entry) == enbl.ENBL_E_PRECONDITION
    zero = np.zeros(n)
    zero[5] = 4.9e-324          # denorm_min counts as zero
    assert _table_call(enbl, gpu_ctx, itype, zero, entry) == enbl.ENBL_E_PRECONDITION
    zero[5] = 1e-300            # one weight above denorm_min is enough
    assert _table_call(enbl, gpu_ctx, itype, zero, entry) == enbl.ENBL_OK
    nan = ok.copy()
    nan[7] = np.nan
    assert _table_call(enbl, gpu_ctx, itype, nan, entry) == enbl.ENBL_OK


@pytest.mark.parametrize("entry", ["mask", "enfuse", "batch"])
def test_exposure_table_for_float_image_unsupported(enbl, gpu_ctx, entry):
    """The reference evaluates the weight function per float pixel
    (src/enfuse.h:637-640); a table cannot stand for it, so a table given for
    a float image is an error, not silently replaced by the built-in."""
    print(f"float image with exposure table entry={entry}", flush=True)
    assert _table_call(enbl, gpu_ctx, F32, np.linspace(0.0, 1.0, 65536), entry) == enbl.ENBL_E_UNSUPPORTED
    assert _table_call(enbl, gpu_ctx, F32, None, entry) == enbl.ENBL_OK


@pytest.mark.parametrize("itype", [U8, U16])
def test_exposure_table_enfuse_matches_oracle(enbl, gpu_ctx, oracle, itype):
    """A plugin's table (a linear ramp, as doc/examples/enfuse/linear.cc's
    weight) through the whole fusion, against the oracle's table path."""
    rng = np.random.default_rng(880 + itype)
    h, w, K = 90, 117, 3
    n = 256 if itype == U8 else 65536
    tab = np.arange(n, dtype=np.float64) / (n - 1)
    imgs = _stack(rng, itype, K, h, w)
    alphas = [np.full((h, w), 255, np.uint8) for _ in range(K)]
    L = oracle.auto_levels(w, h)
    kw = dict(w_exposure=1.0, w_saturation=0.2, exposure_table=tab)
    print(f"enfuse with exposure table itype={itype}", flush=True)
    want, want_a = oracle.enfuse(imgs, alphas, itype, orc_params(**kw), L, dither=1, seed=4)
    dimgs = [torch.from_numpy(i).cuda() for i in imgs]
    das = [torch.from_numpy(a).cuda() for a in alphas]
    out = torch.zeros_like(dimgs[0])
    oa = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    enbl.enfuse(gpu_ctx, itype, dimgs, das, enbl.enfuse_params(**kw), L, out, oa, dither=1, seed=4)
    gpu_ctx.sync()
    np.testing.assert_array_equal(oa.cpu().numpy(), want_a)
    np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_float_contrast_known_answers(enbl, gpu_ctx, oracle):
    """The float local standard deviation's known answers through the GPU
    contrast weight (equal channels, so gray = the sample): a constant 5x5
    window of 4097.0f gives NaN from the float sums, and sliding past a
    column of 8192.0f gives NaN at the next window while the first window
    gives RN(sqrt(11182079)); see tests/test_oracle_enfuse.py for the
    derivations.  Plus a larger image against the oracle's float
    std-dev."""
    rng = np.random.default_rng(12)
    g3 = np.ones((5, 6), np.float32)
    g3[:, 0] = 8192.0
    for g in (np.full((5, 5), 4097.0, np.float32), g3,
              (rng.integers(0, 1 << 12, size=(61, 130)) / 64.0).astype(np.float32)):
        h, w = g.shape
        print(f"float contrast known answer {w}x{h}", flush=True)
        img = np.ascontiguousarray(np.repeat(g[..., None], 3, axis=2))
        a = np.full((h, w), 255, np.uint8)
        kw = dict(w_exposure=0.0, w_contrast=1.0, contrast_window=5)
        dimg, da = torch.from_numpy(img).cuda(), torch.from_numpy(a).cuda()
        dm = torch.full((h, w), -1.0, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        enbl.check(enbl.lib().enbl_enfuse_mask(gpu_ctx.h, F32, w, h, _vp(dimg), 3 * w, _vp(da), w,
                                               ctypes.byref(enbl.enfuse_params(**kw)), _vp(dm), w))
        gpu_ctx.sync()
        got = dm.cpu().numpy()
        np.testing.assert_array_equal(got, oracle.local_stddev_f32(g, a, 5))
        if (h, w) == (5, 5):
            assert np.isnan(got[2, 2])
        elif (h, w) == (5, 6):
            assert got[2, 2] == np.sqrt(np.float32(11182079.0), dtype=np.float32) and np.isnan(got[2, 3])


@pytest.mark.parametrize("itype,K", [(U16, 9), (U8, 3)])
def test_enfuse_streaming_levels_and_shared_tails_match_oracle(enbl, gpu_ctx, oracle, itype, K):
    """2200x2000 images: level 1 takes the fused streaming passes (built on
    the auxiliary stream, beside the weights), the levels below of all
    pyramids go through shared multi-job launches (with 9 images, more jobs
    per level than one launch holds); two calls in a row on one context."""
    h, w = 2000, 2200
    rng = np.random.default_rng(1300 + itype + K)
    yy, xx = np.mgrid[0:h, 0:w]
    base = 0.5 + 0.4 * np.sin(xx / 37.0) * np.cos(yy / 23.0)
    m = 255 if itype == U8 else 65535
    dt = np.uint8 if itype == U8 else np.uint16
    imgs = [np.clip(np.rint(np.clip(base[..., None] * rng.uniform(0.7, 1.3, 3) * 2.0 ** (k - K // 2), 0, 1)
                            ** (1 / 2.2) * m + rng.normal(0, m / 2000.0, (h, w, 3))), 0, m).astype(dt)
            for k in range(K)]
    alphas = [np.where(rng.random((h, w)) < 0.97, 255, 0).astype(np.uint8) for _ in range(K)]
    L = oracle.auto_levels(w, h)
    kw = dict(w_exposure=1.0, w_saturation=0.2, w_contrast=0.3, contrast_window=5)
    want, want_a = oracle.enfuse(imgs, alphas, itype, orc_params(**kw), L, dither=1, seed=2)
    dimgs = [torch.from_numpy(i).cuda() for i in imgs]
    das = [torch.from_numpy(a).cuda() for a in alphas]
    for rep in range(2):
        out = torch.zeros_like(dimgs[0])
        oa = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        print(f"enfuse {w}x{h} itype={itype} K={K} L={L} call {rep}", flush=True)
        enbl.enfuse(gpu_ctx, itype, dimgs, das, enbl.enfuse_params(**kw), L, out, oa, dither=1, seed=2)
        gpu_ctx.sync()
        np.testing.assert_array_equal(oa.cpu().numpy(), want_a)
        np.testing.assert_array_equal(out.cpu().numpy(), want)
