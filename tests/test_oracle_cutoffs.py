"""Pins the oracle's --exposure-cutoff (CutoffExposureFunctor with its
default anti-value / value projectors, src/enfuse.h:654-720) and
--entropy-cutoff (the image clamped before the local entropy,
src/enfuse.h:1145-1188) against the uncut weights and numpy clamps."""
import numpy as np
import pytest

from oracle_binding import F32, U8, U16, enfuse_params, load_oracle

NP = {U8: np.uint8, U16: np.uint16, F32: np.float32}
MAX = {U8: 255, U16: 65535, F32: 1.0}


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


def _img(rng, itype, h, w):
    if itype == F32:
        return rng.uniform(0, 1, size=(h, w, 3)).astype(np.float32)
    return rng.integers(0, MAX[itype] + 1, size=(h, w, 3)).astype(NP[itype])


@pytest.mark.parametrize("itype", [U8, U16, F32])
def test_exposure_cutoff_zeroes_pixels_outside(orc, itype):
    rng = np.random.default_rng(itype)
    h, w = 20, 27
    img = _img(rng, itype, h, w)
    alpha = np.where(rng.random((h, w)) < 0.9, 255, 0).astype(np.uint8)
    lo, hi = (0.2 * MAX[itype], 0.8 * MAX[itype])
    if itype != F32:
        lo, hi = float(int(lo)), float(int(hi))
    base = orc.enfuse_mask(img, alpha, itype, enfuse_params(w_exposure=1.0))
    cut = orc.enfuse_mask(img, alpha, itype, enfuse_params(w_exposure=1.0, exposure_cutoff=(lo, hi)))
    keep = (img.min(axis=2).astype(np.float64) >= lo) & (img.max(axis=2).astype(np.float64) <= hi)
    assert 0 < keep.sum() < h * w
    np.testing.assert_array_equal(cut, np.where(keep, base, 0).astype(np.float32))
    # cutoffs at the ends of the range cut nothing
    full = orc.enfuse_mask(img, alpha, itype, enfuse_params(w_exposure=1.0, exposure_cutoff=(0.0, float(MAX[itype]))))
    np.testing.assert_array_equal(full, base)
    with pytest.raises(ValueError):
        orc.enfuse_mask(img, alpha, itype, enfuse_params(w_exposure=1.0, exposure_cutoff=(hi, lo)))


def _clamp(img, lo, up, mx):
    r, g, b = img[..., 0], img[..., 1], img[..., 2]
    out = np.empty_like(img)
    out[..., 0] = np.where(r <= lo, 0, np.where(r >= up, mx, r))
    out[..., 1] = np.where(g <= lo, 0, np.where(g >= up, mx, g))
    out[..., 2] = np.where(r <= lo, 0, np.where(b >= up, mx, b))   # blue tests red, as the reference does
    return out


@pytest.mark.parametrize("itype", [U8, U16, F32])
def test_entropy_cutoff_is_entropy_of_the_clamped_image(orc, itype):
    rng = np.random.default_rng(10 + itype)
    h, w = 22, 25
    img = _img(rng, itype, h, w)
    alpha = np.where(rng.random((h, w)) < 0.9, 255, 0).astype(np.uint8)
    lo, up = 0.3 * MAX[itype], 0.7 * MAX[itype]
    if itype != F32:
        lo, up = float(int(lo)), float(int(up))
    prm = dict(w_exposure=0.0, w_entropy=1.0, entropy_window=3)
    cut = orc.enfuse_mask(img, alpha, itype, enfuse_params(**prm, entropy_cutoff=(lo, up)))
    t = NP[itype]
    clamped = _clamp(img, t(lo), t(up), t(MAX[itype]) if itype != F32 else np.float32(np.finfo(np.float32).max))
    want = orc.enfuse_mask(clamped, alpha, itype, enfuse_params(**prm))
    np.testing.assert_array_equal(cut, want)
    # the blue-channel quirk is exercised
    assert ((img[..., 0] <= t(lo)) & (img[..., 2] > t(lo))).any()


def test_entropy_cutoff_full_range_is_identity_for_integers(orc):
    rng = np.random.default_rng(3)
    img = _img(rng, U8, 18, 21)
    alpha = np.full((18, 21), 255, np.uint8)
    prm = dict(w_exposure=0.0, w_entropy=1.0, entropy_window=5)
    a = orc.enfuse_mask(img, alpha, U8, enfuse_params(**prm))
    b = orc.enfuse_mask(img, alpha, U8, enfuse_params(**prm, entropy_cutoff=(0.0, 255.0)))
    np.testing.assert_array_equal(a, b)
