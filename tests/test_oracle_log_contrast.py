"""Pins the oracle's contrast-by-edges mode (Laplacian of Gaussian, with
optional local contrast enhancement; src/enfuse.h:1020-1093).

VIGRA's kernels and convolution are not vendored, so the pins are what the
mathematics fixes: the kernels' normalisations and shapes, the LoG of
constant, linear and quadratic images, the mirror border against numpy's own
`reflect` padding, and the reference's own clamping and fill-in functors."""
import numpy as np
import pytest

from oracle_binding import F32, U8, U16, enfuse_params, load_oracle


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


@pytest.mark.parametrize("sigma", [0.3, 0.5, 0.8, 1.0, 1.7, 3.2])
def test_kernel_normalisations_and_shapes(orc, sigma):
    s = orc.vigra_kernel(sigma, 0)
    d = orc.vigra_kernel(sigma, 2)
    rs, rd = len(s) // 2, len(d) // 2
    assert rs == max(1, int(3.0 * sigma + 0.5)) and rd == max(1, int(4.0 * sigma + 0.5))
    x = np.arange(-rs, rs + 1, dtype=np.float64)
    xd = np.arange(-rd, rd + 1, dtype=np.float64)
    # symmetric bit for bit, smoothing sums to 1, Gaussian shape
    np.testing.assert_array_equal(s, s[::-1])
    np.testing.assert_array_equal(d, d[::-1])
    assert abs(s.sum() - 1.0) < 1e-15
    np.testing.assert_allclose(s / s[rs], np.exp(-x * x / (2 * sigma * sigma)), rtol=1e-13)
    # second derivative: no DC, sum k x^2 / 2 = 1, and k + dc has the shape
    # (1 - (x/sigma)^2) exp(-x^2 / 2 sigma^2)
    assert abs(d.sum()) < 1e-14
    assert abs((d * xd * xd).sum() / 2.0 - 1.0) < 1e-14
    shape = (1 - (xd / sigma) ** 2) * np.exp(-xd * xd / (2 * sigma * sigma))
    dc = shape.mean()
    np.testing.assert_allclose(d / d[rd], (shape - dc) / (shape[rd] - dc), rtol=1e-12, atol=1e-14)


def _mask(orc, img, itype, **kw):
    h, w, _ = img.shape
    alpha = np.full((h, w), 255, np.uint8)
    prm = enfuse_params(w_exposure=0.0, w_contrast=1.0, **kw)
    return orc.enfuse_mask(img, alpha, itype, prm)


def _rgb(gray, dtype):
    return np.repeat(gray[..., None], 3, axis=2).astype(dtype)


@pytest.mark.parametrize("sigma", [0.5, 1.0, 1.6])
def test_log_of_polynomials(orc, sigma):
    h, w = 40, 48
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    r = len(orc.vigra_kernel(sigma, 2)) // 2
    inner = (slice(r, h - r), slice(r, w - r))
    # constant and linear images: no curvature
    for g in (np.full((h, w), 77.0), 3.0 * xx + 2.0 * yy + 5.0):
        m = _mask(orc, _rgb(g, np.float32), F32, edge_scale=sigma)
        assert np.abs(m[inner]).max() < 1e-9
    # x^2 + y^2: the Laplacian is 4 everywhere the window stays inside, and
    # mirroring at x = 0, y = 0 keeps the quadratic, so the top-left border
    # too (float images keep their intermediate images in float: values up to
    # 3900 leave about 1e-4 of rounding)
    g = xx * xx + yy * yy
    m = _mask(orc, _rgb(g, np.float32), F32, edge_scale=sigma)
    np.testing.assert_allclose(m[: h - r, : w - r], 4.0, rtol=1e-4)
    # 16-bit images keep them in double: the rounded magnitude is exactly 4
    m = _mask(orc, _rgb(g, np.uint16), U16, edge_scale=sigma)
    assert (m[: h - r, : w - r] == np.float32(4.0 / 65535.0)).all()
    # an 8-bit image of x^2 + y^2 (at most 200): the rounded magnitude is 4,
    # the weight 4 / 255
    n8 = 11
    y8, x8 = np.mgrid[0:n8, 0:n8]
    m = _mask(orc, _rgb(x8 * x8 + y8 * y8, np.uint8), U8, edge_scale=sigma)
    assert (m[: n8 - r, : n8 - r] == np.float32(4.0 / 255.0)).all()


def _conv_reflect(a, k, axis):
    r = len(k) // 2
    pad = [(0, 0), (0, 0)]
    pad[axis] = (r, r)
    p = np.pad(a, pad, mode="reflect")
    out = np.zeros_like(a, dtype=np.float64)
    n = a.shape[axis]
    for t in range(2 * r + 1):
        sl = [slice(None), slice(None)]
        sl[axis] = slice(t, t + n)
        out += k[t] * p[tuple(sl)]
    return out


@pytest.mark.parametrize("sigma,lce", [(0.7, None), (1.3, None), (0.9, (2.5, 0.6))])
def test_log_border_and_lce_against_numpy(orc, sigma, lce):
    """Float images: the oracle's LoG magnitude against numpy's reflect-padded
    separable convolution with the same kernels (mirror without repeating the
    edge is numpy's 'reflect'), to float rounding."""
    rng = np.random.default_rng(int(sigma * 10))
    h, w = 23, 31
    g = rng.uniform(0, 1, size=(h, w)).astype(np.float32)
    img = _rgb(g, np.float32)
    gray = ((img[..., 0] + img[..., 1] + img[..., 2]).astype(np.float64) / 3.0).astype(np.float32).astype(np.float64)
    kw = {}
    src = gray
    if lce:
        ls, f = lce
        kl = orc.vigra_kernel(ls, 0)
        sm = _conv_reflect(_conv_reflect(gray, kl, 1), kl, 0)
        src = (1.0 + f) * gray - f * sm
        kw = {"lce_scale": ls, "lce_factor": f}
    s = orc.vigra_kernel(sigma, 0)
    d = orc.vigra_kernel(sigma, 2)
    lap = _conv_reflect(_conv_reflect(src, d, 1), s, 0) + _conv_reflect(_conv_reflect(src, s, 1), d, 0)
    m = _mask(orc, img, F32, edge_scale=sigma, **kw)
    np.testing.assert_allclose(m, np.abs(lap), rtol=2e-5, atol=2e-6)


def test_clamping_and_fill_in(orc):
    """min_curvature <= 0: magnitudes at or below -min_curvature become 0
    (ClampingFunctor); > 0: magnitudes below it are replaced by the local
    standard deviation scaled by min_curvature / 255 (FillInFunctor)."""
    rng = np.random.default_rng(5)
    h, w = 30, 34
    g = rng.integers(0, 256, size=(h, w)).astype(np.uint8)
    img = _rgb(g, np.uint8)
    base = _mask(orc, img, U8, edge_scale=0.8)
    lap = np.rint(base * 255.0)   # weight = magnitude / 255 exactly for small integers
    np.testing.assert_allclose(base, (lap / 255.0).astype(np.float32), rtol=0, atol=0)
    clamped = _mask(orc, img, U8, edge_scale=0.8, min_curvature=-20.0)
    np.testing.assert_array_equal(clamped, np.where(lap <= 20, 0, lap / 255.0).astype(np.float32))
    sd = _mask(orc, img, U8, contrast_window=5)          # the local standard deviation mode
    sdv = np.rint(sd * 255.0)
    filled = _mask(orc, img, U8, edge_scale=0.8, min_curvature=30.0, contrast_window=5)
    fill = np.floor(30.0 / 255.0 * sdv + 0.5)
    expect = np.where(lap >= 30, lap, fill) / 255.0
    np.testing.assert_array_equal(filled, expect.astype(np.float32))


def test_kernel_longer_than_line(orc):
    img = np.zeros((6, 40, 3), np.uint16)
    with pytest.raises(ValueError):
        _mask(orc, img, U16, edge_scale=2.0)     # derivative radius 8 >= 6 rows
