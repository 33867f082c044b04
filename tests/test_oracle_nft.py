"""Pins of the oracle's nearest-feature-transform seam mask and its float
Felzenszwalb-Huttenlocher distance transform (oracle/oracle_nft.cpp).

Where every value the transform forms is an integer below 2^24 (rows and
columns shorter than 4096) its float arithmetic is exact, so it must equal
the brute-force distance to the nearest feature pixel; the Manhattan
transform is exact at every size.  Longer rows round q^2 and the envelope
intersections in float: there every output is checked per pixel to be one
of the reference's float candidates within the intersection-rounding bound
of the best, and the operation order is pinned by a hand-derived known
answer that a reordered intersection formula fails; the GPU is compared
with the oracle bit for bit."""
import numpy as np
import pytest

FLT_MAX = np.float32(np.finfo(np.float32).max)


def _brute_d2(src):
    """Exact squared Euclidean distance to the nearest non-zero pixel (int64;
    -1 where there is none)."""
    h, w = src.shape
    ys, xs = np.nonzero(src)
    if len(ys) == 0:
        return np.full((h, w), -1, np.int64)
    yy, xx = np.mgrid[0:h, 0:w]
    best = np.full((h, w), np.iinfo(np.int64).max, np.int64)
    for y, x in zip(ys, xs):
        best = np.minimum(best, (yy - y) ** 2 + (xx - x) ** 2)
    return best


def _brute_l1(src):
    h, w = src.shape
    ys, xs = np.nonzero(src)
    yy, xx = np.mgrid[0:h, 0:w]
    best = np.full((h, w), np.iinfo(np.int64).max, np.int64)
    for y, x in zip(ys, xs):
        best = np.minimum(best, np.abs(yy - y) + np.abs(xx - x))
    return best


def _random_binary(rng, h, w, density):
    return (rng.random((h, w)) < density).astype(np.uint8) * rng.integers(1, 256)


SHAPES = [(1, 1), (1, 9), (9, 1), (2, 2), (7, 5), (16, 33), (40, 27), (64, 64)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("density", [0.02, 0.2, 0.7])
def test_euclidean_transform_equals_brute_force(oracle, shape, density):
    rng = np.random.default_rng(hash((shape, density)) & 0xffff)
    h, w = shape
    src = _random_binary(rng, h, w, density)
    got = oracle.distance_transform(src, 2)
    d2 = _brute_d2(src)
    if (d2 < 0).all():
        np.testing.assert_array_equal(got, np.full((h, w), np.sqrt(FLT_MAX), np.float32))
    else:
        np.testing.assert_array_equal(got, np.sqrt(d2.astype(np.float32)))


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("density", [0.02, 0.3])
def test_manhattan_transform_equals_brute_force(oracle, shape, density):
    rng = np.random.default_rng(7 + (hash((shape, density)) & 0xffff))
    h, w = shape
    src = _random_binary(rng, h, w, density)
    got = oracle.distance_transform(src, 1)
    if not src.any():
        np.testing.assert_array_equal(got, np.full((h, w), FLT_MAX, np.float32))
    else:
        np.testing.assert_array_equal(got, _brute_l1(src).astype(np.float32))


def test_transform_known_answers(oracle):
    src = np.zeros((5, 7), np.uint8)
    src[2, 3] = 1
    want = np.sqrt(((np.mgrid[0:5, 0:7][0] - 2) ** 2 + (np.mgrid[0:5, 0:7][1] - 3) ** 2).astype(np.float32))
    np.testing.assert_array_equal(oracle.distance_transform(src, 2), want)
    # features along one column only: every column without features takes
    # its distance from the row pass
    src = np.zeros((6, 40), np.uint8)
    src[:, 0] = 1
    np.testing.assert_array_equal(oracle.distance_transform(src, 2), np.tile(np.arange(40, dtype=np.float32), (6, 1)))


def _column_pass(src):
    """Exact squared distance to the nearest feature of the same column, as
    float32 (exact below 2^24), FLT_MAX where the column has none."""
    h, w = src.shape
    ys = np.arange(h)
    f = np.full((h, w), FLT_MAX, np.float32)
    for x in np.nonzero(src.any(axis=0))[0]:
        fy = np.nonzero(src[:, x])[0]
        f[:, x] = np.min((ys[:, None] - fy[None, :]) ** 2, axis=1).astype(np.float32)
    return f


def _row_candidates(frow):
    """For every sample q of a row: the best and second-best float candidate
    RN(RN((q - v)^2) + f[v]) over all v with a finite f[v] (the values the
    reference's output sweep can produce, src/openmp_vigra.h:341-349), and
    a membership test for a given output."""
    w = frow.size
    fin = np.nonzero(frow < FLT_MAX)[0]
    q = np.arange(w)
    best = np.full(w, FLT_MAX)
    second = np.full(w, FLT_MAX)
    chunks = []
    for c0 in range(0, fin.size, 256):
        v = fin[c0:c0 + 256]
        cand = ((q[:, None] - v[None, :]) ** 2).astype(np.float32) + frow[v][None, :]
        chunks.append(cand)
        allc = np.sort(np.concatenate([best[:, None], second[:, None], cand], axis=1), axis=1)
        best, second = allc[:, 0], allc[:, 1]
    return best, second, chunks


@pytest.mark.parametrize("kind", ["dense", "sparse", "far"])
def test_long_rows_per_pixel(oracle, kind):
    """Rows of 7168 samples (the headline canvas width), per pixel.  The row
    pass input is the exact column pass; every output of the row transform
    must be one of the float candidates RN(RN((q - v)^2) + f[v]), never
    below the best one, equal to it wherever it is below 2^24 and no sum the
    envelope forms rounds (q < 4096), and otherwise within the error the
    rounded envelope intersections allow: a sample assigned to the wrong one
    of two neighbouring parabolas lies within |ds| <= 10 / (2 dq) of their
    true intersection (sums below 2^26 round by at most 4 each, the
    difference by at most 2), so its value exceeds the best by at most
    2 dq |ds| <= 10, plus the rounding of the candidate itself.  (The
    reference's float envelope does pick the second-best candidate at a few
    pixels of such rows, so exact equality everywhere would be wrong.)  The
    whole transform is the square root of this row pass."""
    rng = np.random.default_rng({"dense": 1, "sparse": 2, "far": 3}[kind])
    h, w = (12, 7168) if kind == "dense" else (24, 7168)
    src = np.zeros((h, w), np.uint8)
    if kind == "dense":
        src[rng.random((h, w)) < 0.05] = 1
    elif kind == "sparse":
        src[rng.integers(0, h, 7), rng.integers(0, w, 7)] = 1
    else:
        src[0, rng.integers(0, w, 300)] = 1          # column distances up to (h-1)^2, every row
        src[:, :64] = 0
    f = _column_pass(src)
    full = oracle.distance_transform(src, 2)
    picked_second = 0
    for y in range(h):
        out = oracle.euclid_1d(f[y])
        np.testing.assert_array_equal(full[y], np.sqrt(out, dtype=np.float32), err_msg=f"row {y}")
        best, second, chunks = _row_candidates(f[y])
        member = np.zeros(w, bool)
        for cand in chunks:
            member |= np.any(cand == out[:, None], axis=1)
        assert member.all(), f"row {y}: output not a candidate at {np.nonzero(~member)[0][:5]}"
        assert np.all(out >= best)
        exact = (np.arange(w) < 4096) & (best < 2.0 ** 24)
        np.testing.assert_array_equal(out[exact], best[exact], err_msg=f"row {y}")
        ulp = (np.nextafter(best, np.float32(np.inf)) - best).astype(np.float64)
        assert np.all(out.astype(np.float64) - best <= 10.0 + 2 * ulp), f"row {y}"
        picked_second += int(np.count_nonzero(out != best))
    if kind == "dense":
        assert picked_second > 0      # the rounding regime is exercised


def test_envelope_known_answer_rounding_order(oracle):
    """Four samples f = (18839148, 18839132, 18839148, 18839126), by hand in
    the reference's float order (src/openmp_vigra.h:320-349).  q = 1:
    sum = RN(18839133) = 18839132, s = (18839132 - 18839148) / 2 = -8.
    q = 2: sum = 18839152, f[1] + 1 rounds to 18839132, s = 20 / 2 = 10.
    q = 3: sum = RN(18839135) = 18839136 (tie to even); against v = 2,
    s = (18839136 - 18839152) / 2 = -8 <= 10 pops it; against v = 1,
    s = (18839136 - 18839132) / 4 = 1 > -8, so z = (-inf, -8, 1).  The sweep
    keeps sample 1 on v = 1 (z[2] = 1 is not < 1): d = (18839132,
    18839132, RN(18839127) = 18839128, 18839126), although v = 3 gives
    18839130 at sample 1 (true intersection 0.5).  The reordered formula
    ((sum_q - f[v]) - v^2) / (2 dq) gives s = 3 / 4 at q = 3 and assigns
    sample 1 to v = 3, so it fails this answer."""
    f = np.array([18839148, 18839132, 18839148, 18839126], np.float32)
    want = np.array([18839132, 18839132, 18839128, 18839126], np.float32)
    np.testing.assert_array_equal(oracle.euclid_1d(f), want)
    planted = oracle.euclid_1d(f, planted=1)
    assert planted[1] == np.float32(18839130)
    # the same four samples embedded in a longer row at the same offsets
    # keep the answer (the envelope depends only on these parabolas)
    g = np.full(9, FLT_MAX, np.float32)
    g[:4] = f
    np.testing.assert_array_equal(oracle.euclid_1d(g)[:4], want)


def _blob_alphas(rng, h, w):
    yy, xx = np.mgrid[0:h, 0:w]
    cx, cy = 0.3 * w + rng.normal(0, 1), 0.5 * h
    white = (((xx - cx) / (0.45 * w)) ** 2 + ((yy - cy) / (0.48 * h)) ** 2 < 1).astype(np.uint8) * 255
    black = (((xx - 0.7 * w) / (0.45 * w)) ** 2 + ((yy - cy) / (0.48 * h)) ** 2 < 1).astype(np.uint8) * 255
    return white, black


@pytest.mark.parametrize("shape", [(24, 40), (37, 61), (64, 48)])
def test_nft_mask_equals_brute_force(oracle, shape):
    rng = np.random.default_rng(shape[0] * 100 + shape[1])
    h, w = shape
    white, black = _blob_alphas(rng, h, w)
    mask, rc = oracle.nft_mask(white, black)
    assert rc == 0
    d12 = _brute_d2((white > black).astype(np.uint8))
    d21 = _brute_d2((black > white).astype(np.uint8))
    np.testing.assert_array_equal(mask, np.where(d12 < d21, 255, 0).astype(np.uint8))


def _edge_alphas(rng, h, w):
    """White around the left edge, black around the right edge and a band
    across the top: the strips make them neighbours across the borders."""
    yy, xx = np.mgrid[0:h, 0:w]
    white = ((xx < 0.35 * w + rng.integers(0, 3)) & (yy > 0.2 * h)).astype(np.uint8) * 255
    black = (((xx > 0.6 * w) & (yy > 0.1 * h)) | (yy < 0.08 * h + 1)).astype(np.uint8) * 255
    return white, black


@pytest.mark.parametrize("shape", [(24, 40), (37, 61), (30, 31)])
@pytest.mark.parametrize("boundary", [1, 2, 3])
def test_periodic_nft_mask_equals_brute_force(oracle, shape, boundary):
    """Strip boundaries: the brute-force transform of the difference images
    tiled twice along each wrapped axis, cut back by taking position p + n
    for p < n / 2 on a doubled axis of length n."""
    rng = np.random.default_rng(boundary * 1000 + shape[0] * 10 + shape[1])
    h, w = shape
    white, black = _edge_alphas(rng, h, w)
    mask, rc = oracle.nft_mask_periodic(white, black, boundary)
    ty, tx = (2 if boundary in (2, 3) else 1), (2 if boundary in (1, 3) else 1)
    d12 = _brute_d2(np.tile((white > black).astype(np.uint8), (ty, tx)))
    d21 = _brute_d2(np.tile((black > white).astype(np.uint8), (ty, tx)))
    ys = np.arange(h) + (h if ty == 2 else 0) * (np.arange(h) < h // 2)
    xs = np.arange(w) + (w if tx == 2 else 0) * (np.arange(w) < w // 2)
    want = np.where(d12[np.ix_(ys, xs)] < d21[np.ix_(ys, xs)], 255, 0).astype(np.uint8)
    np.testing.assert_array_equal(mask, want)
    # the wrap matters here: the open-boundary mask differs
    if boundary == 1:
        assert (oracle.nft_mask(white, black)[0] != mask).any()


def test_nft_mask_ties_go_to_black(oracle):
    """Equidistant pixels compare dist12 < dist21 false: 0 (the black image),
    src/nearest.h:540-545; a planted <= gives a different mask."""
    h, w = 40, 31
    white = np.zeros((h, w), np.uint8)
    black = np.zeros((h, w), np.uint8)
    white[:, :12] = 255
    black[:, 19:] = 255
    white[:, 12:19] = black[:, 12:19] = 200   # common, equal alphas
    mask, rc = oracle.nft_mask(white, black)
    assert rc == 0
    want = np.zeros((h, w), np.uint8)
    want[:, :15] = 255                        # x = 15 is 4 from both sides
    np.testing.assert_array_equal(mask, want)
    d12 = oracle.distance_transform((white > black).astype(np.uint8))
    d21 = oracle.distance_transform((black > white).astype(np.uint8))
    np.testing.assert_array_equal(mask, np.where(d12 < d21, 255, 0).astype(np.uint8))
    assert not np.array_equal(mask, np.where(d12 <= d21, 255, 0).astype(np.uint8))


def test_nft_mask_difference_is_saturating(oracle):
    """Partial alphas: the features of the white image are where its alpha
    exceeds the black alpha (saturating difference), overlap included; a
    planted "white set and black unset" rule differs."""
    rng = np.random.default_rng(5)
    h, w = 30, 50
    white, black = _blob_alphas(rng, h, w)
    white = np.where(white > 0, 180, 0).astype(np.uint8)   # the overlap belongs to white
    black = np.where(black > 0, 120, 0).astype(np.uint8)
    mask, rc = oracle.nft_mask(white, black)
    assert rc == 0
    d12 = _brute_d2((white > black).astype(np.uint8))
    d21 = _brute_d2((black > white).astype(np.uint8))
    np.testing.assert_array_equal(mask, np.where(d12 < d21, 255, 0).astype(np.uint8))
    planted = np.where(_brute_d2(((white != 0) & (black == 0)).astype(np.uint8)) < d21, 255, 0)
    assert not np.array_equal(mask, planted)


def test_nft_excessive_overlap(oracle):
    """Fewer than 2 * 2 * (w + h) exclusive pixels in both images is the
    reference's excessive-overlap exit (src/nearest.h:449-457, :523-538)."""
    h, w = 20, 30
    threshold = 4 * (w + h)
    white = np.full((h, w), 255, np.uint8)
    black = white.copy()
    assert oracle.nft_mask(white, black)[1] == -2
    flat = black.reshape(-1)
    flat[: threshold - 1] = 0                # white-only pixels: threshold - 1
    assert oracle.nft_mask(white, black)[1] == -2
    flat[: threshold] = 0
    assert oracle.nft_mask(white, black)[1] == 0
