"""GPU nearest-feature-transform seam mask and float distance transform
(nft.hip) against the oracle, bit for bit, including columns and rows
longer than 4096 where the float roundings of the reference decide the
result."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _gpu_dt(enbl, ctx, src, norm):
    h, w = src.shape
    pitch = w + 3
    d_src = torch.zeros((h, pitch), dtype=torch.uint8, device="cuda")
    d_src[:, :w] = torch.from_numpy(np.ascontiguousarray(src)).cuda()
    d_out = torch.full((h, w + 1), -1.0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    enbl.check(enbl.lib().enbl_distance_transform(ctx.h, w, h, _vp(d_src), pitch, norm, _vp(d_out), w + 1))
    ctx.sync()
    out = d_out.cpu().numpy()
    assert (out[:, w] == -1.0).all(), "wrote past the row"
    return out[:, :w]


def _binary(rng, h, w, density):
    return ((rng.random((h, w)) < density) * rng.integers(1, 256, size=(h, w))).astype(np.uint8)


@pytest.mark.parametrize("shape", [(1, 1), (1, 300), (300, 1), (64, 65), (257, 513), (600, 4500), (4200, 300)])
@pytest.mark.parametrize("norm", [1, 2])
@pytest.mark.parametrize("density", [0.0005, 0.05, 0.6])
def test_distance_transform_matches_oracle(enbl, gpu_ctx, oracle, shape, norm, density):
    h, w = shape
    print(f"distance transform {w}x{h} norm={norm} density={density}", flush=True)
    rng = np.random.default_rng(h * 7919 + w * 31 + norm + int(density * 1e4))
    src = _binary(rng, h, w, density)
    np.testing.assert_array_equal(_gpu_dt(enbl, gpu_ctx, src, norm), oracle.distance_transform(src, norm))


@pytest.mark.parametrize("norm", [1, 2])
def test_distance_transform_without_features(enbl, gpu_ctx, oracle, norm):
    src = np.zeros((37, 53), np.uint8)
    np.testing.assert_array_equal(_gpu_dt(enbl, gpu_ctx, src, norm), oracle.distance_transform(src, norm))


def _polygon_alphas(rng, h, w):
    """Two overlapping convex regions (ellipses with a random tilt)."""
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    out = []
    for cx in (0.32, 0.68):
        th = rng.uniform(-0.3, 0.3)
        u = (xx - cx * w) * np.cos(th) + (yy - 0.5 * h) * np.sin(th)
        v = -(xx - cx * w) * np.sin(th) + (yy - 0.5 * h) * np.cos(th)
        out.append(((u / (0.4 * w)) ** 2 + (v / (0.47 * h)) ** 2 < 1).astype(np.uint8) * 255)
    return out


def _gpu_nft(enbl, ctx, white, black, norm=2):
    h, w = white.shape
    ap, mp = w + 5, w + 2
    dw = torch.zeros((h, ap), dtype=torch.uint8, device="cuda")
    db = torch.zeros((h, ap), dtype=torch.uint8, device="cuda")
    dw[:, :w] = torch.from_numpy(white).cuda()
    db[:, :w] = torch.from_numpy(black).cuda()
    dm = torch.full((h, mp), 7, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    rc = enbl.lib().enbl_nft_mask(ctx.h, w, h, _vp(dw), ap, _vp(db), ap, norm, _vp(dm), mp)
    ctx.sync()
    m = dm.cpu().numpy()
    assert (m[:, w:] == 7).all(), "wrote past the row"
    return m[:, :w], rc


@pytest.mark.parametrize("shape", [(96, 160), (1536, 2048), (700, 5000)])
@pytest.mark.parametrize("norm", [1, 2])
def test_nft_mask_matches_oracle(enbl, gpu_ctx, oracle, shape, norm):
    h, w = shape
    print(f"nft mask {w}x{h} norm={norm}", flush=True)
    rng = np.random.default_rng(h + w + norm)
    white, black = _polygon_alphas(rng, h, w)
    want, want_rc = oracle.nft_mask(white, black, norm)
    got, rc = _gpu_nft(enbl, gpu_ctx, white, black, norm)
    assert want_rc == 0 and rc == 0
    np.testing.assert_array_equal(got, want)


def test_nft_mask_excessive_overlap(enbl, gpu_ctx, oracle):
    """Identical alphas: the reference's excessive-overlap exit becomes
    ENBL_E_PRECONDITION; the mask is still written and equals the oracle's."""
    h, w = 40, 60
    white = np.full((h, w), 255, np.uint8)
    black = white.copy()
    black[:5, :5] = 0
    want, want_rc = oracle.nft_mask(white, black)
    got, rc = _gpu_nft(enbl, gpu_ctx, white, black)
    assert want_rc == -2 and rc == enbl.ENBL_E_PRECONDITION
    np.testing.assert_array_equal(got, want)


def _edge_alphas(rng, h, w):
    yy, xx = np.mgrid[0:h, 0:w]
    white = ((xx < 0.35 * w + rng.integers(0, 3)) & (yy > 0.2 * h)).astype(np.uint8) * 255
    black = (((xx > 0.6 * w) & (yy > 0.1 * h)) | (yy < 0.08 * h + 1)).astype(np.uint8) * 255
    return white, black


@pytest.mark.parametrize("shape", [(37, 61), (300, 517), (1200, 2100), (2100, 700)])
@pytest.mark.parametrize("boundary", [1, 2, 3])
def test_periodic_nft_mask_matches_oracle(enbl, gpu_ctx, oracle, shape, boundary):
    h, w = shape
    print(f"periodic nft {w}x{h} boundary={boundary}", flush=True)
    rng = np.random.default_rng(h + 7 * w + boundary)
    white, black = _edge_alphas(rng, h, w)
    want, wrc = oracle.nft_mask_periodic(white, black, boundary)
    dw, db = torch.from_numpy(white).cuda(), torch.from_numpy(black).cuda()
    dm = torch.full((h, w + 2), 7, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    rc = enbl.lib().enbl_nft_mask_periodic(gpu_ctx.h, w, h, _vp(dw), w, _vp(db), w, 2, boundary, _vp(dm), w + 2)
    gpu_ctx.sync()
    assert (rc == enbl.ENBL_E_PRECONDITION) == (wrc == -2)
    got = dm.cpu().numpy()
    assert (got[:, w:] == 7).all()
    np.testing.assert_array_equal(got[:, :w], want)


@pytest.mark.parametrize("w", [2000, 4700])
def test_distance_transform_deep_row_envelopes(enbl, gpu_ctx, oracle, w):
    """Rows whose column distances grow like sqrt(x) keep nearly every sample
    on the lower envelope (stacks thousands deep, far beyond the LDS ring),
    then features in the row itself pop them all: the ring's reloads from
    device memory decide every result."""
    h = 300
    src = np.zeros((h, w), np.uint8)
    xs = np.arange(w)
    d = np.minimum(np.rint(np.sqrt(xs * (h - 1) ** 2 / w)).astype(int), h - 1)
    cut = w - w // 20
    src[d[:cut], xs[:cut]] = 200
    src[0, cut::7] = 255
    src[h // 2, w // 3] = 9
    print(f"deep envelopes {w}x{h}", flush=True)
    np.testing.assert_array_equal(_gpu_dt(enbl, gpu_ctx, src, 2), oracle.distance_transform(src, 2))


def _quotients(enbl, ctx, num, den):
    dn = torch.from_numpy(np.ascontiguousarray(num, np.float32)).cuda()
    dd = torch.from_numpy(np.ascontiguousarray(den, np.int32)).cuda()
    out = torch.empty_like(dn)
    torch.cuda.synchronize()
    enbl.check(enbl.lib().enbl_nft_quotients(ctx.h, _vp(dn), _vp(dd), _vp(out), dn.numel()))
    ctx.sync()
    return out.cpu().numpy()


def test_envelope_reciprocals_are_correctly_rounded(enbl, gpu_ctx):
    """The row pass's quotient is one FMA correction of n * y (Markstein),
    which is the correctly rounded n / d when y is the correctly rounded
    1 / d.  With n = 1 the kernel returns exactly that y-based quotient;
    checked for every integer d below 2^24, i.e. every float significand
    (1 / (2^k d) only shifts the exponent, so larger d are covered)."""
    print("envelope reciprocals d < 2^24", flush=True)
    den = np.arange(1, 1 << 24, dtype=np.int32)
    num = np.ones(den.shape, np.float32)
    got = _quotients(enbl, gpu_ctx, num, den)
    want = np.float32(1.0) / den.astype(np.float32)
    bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
    assert bad.size == 0, f"{bad.size} reciprocals differ, first d={den[bad[:5]]}"


def test_envelope_quotients_equal_ieee_division(enbl, gpu_ctx):
    """(sum_q - sum_v) / (2 (q - v)) on the device equals the IEEE float
    quotient bit for bit, for every even divisor up to 2^16 (rows up to
    32768 samples apart) and 2^17 + 2 and powers of two to 2^30, over
    numerators of every magnitude and both signs (integers, half-integers
    and random significands up to FLT_MAX)."""
    rng = np.random.default_rng(2024)
    nums = [np.float32(v) for v in (1.0, -1.0, 3.0, 0.5, -0.5, 2.5, 16777215.0, 16777217.0,
                                    np.finfo(np.float32).max, -np.finfo(np.float32).max, 2.0 ** -23)]
    # the numerators are differences of sums of at least 1 (multiples of
    # 2^-23 or zero), so quotients stay clear of the subnormal range
    mags = rng.uniform(-6, 38, size=300)
    sig = rng.uniform(1.0, 2.0, size=300) * rng.choice([-1.0, 1.0], size=300)
    nums += list((sig * 10.0 ** mags).astype(np.float32))
    nums += list(rng.integers(-(1 << 26), 1 << 26, size=100).astype(np.float32))
    nums += list((rng.integers(-(1 << 24), 1 << 24, size=100) + 0.5).astype(np.float32))
    num_set = np.array([v for v in nums if np.isfinite(v)], np.float32)
    dens = np.concatenate([np.arange(2, (1 << 16) + 1, 2), [(1 << 17) + 2], 2 ** np.arange(17, 31)]).astype(np.int32)
    print(f"envelope quotients: {num_set.size} numerators x {dens.size} divisors", flush=True)
    num = np.repeat(num_set, dens.size)
    den = np.tile(dens, num_set.size)
    got = _quotients(enbl, gpu_ctx, num, den)
    with np.errstate(over="ignore", under="ignore"):
        want = num / den.astype(np.float32)
    bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
    assert bad.size == 0, f"{bad.size} quotients differ, first {list(zip(num[bad[:5]], den[bad[:5]]))}"
