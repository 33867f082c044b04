"""Pins the oracle's grayscale projectors (--gray-projector, src/mga.h:60-250)
through the exposure weight with an identity table (the weight of a pixel is
its gray code), against closed forms: channel maximum and minimum, rounded
mean of the extreme channels, the rounded luminance and channel-mixer sums,
and CIE L* (116 Y^(1/3) - 16 above 216/24389, 24389/27 Y below) of the sRGB
luminance, with and without the gamma 1/0.45."""
import numpy as np
import pytest

from oracle_binding import (GRAY_ANTI_VALUE, GRAY_LIGHTNESS, GRAY_LSTAR, GRAY_LUMINANCE, GRAY_MIXER,
                            GRAY_PLSTAR, GRAY_VALUE, U8, U16, enfuse_params, load_oracle)


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


def _codes(orc, img, itype, proj, mixer=(0.0, 0.0, 0.0)):
    h, w, _ = img.shape
    n = 256 if itype == U8 else 65536
    prm = enfuse_params(w_exposure=1.0, exposure_table=np.arange(n, dtype=np.float64), gray_projector=proj,
                        mixer=mixer)
    return orc.enfuse_mask(img, np.full((h, w), 255, np.uint8), itype, prm).astype(np.int64)


def _round(v):
    return np.floor(v + 0.5).astype(np.int64)


@pytest.mark.parametrize("itype", [U8, U16])
def test_simple_projectors(orc, itype):
    rng = np.random.default_rng(itype)
    mx = 255 if itype == U8 else 65535
    img = rng.integers(0, mx + 1, size=(13, 17, 3)).astype(np.uint8 if itype == U8 else np.uint16)
    f = img.astype(np.int64)
    hi, lo = f.max(axis=2), f.min(axis=2)
    np.testing.assert_array_equal(_codes(orc, img, itype, GRAY_VALUE), hi)
    np.testing.assert_array_equal(_codes(orc, img, itype, GRAY_ANTI_VALUE), lo)
    np.testing.assert_array_equal(_codes(orc, img, itype, GRAY_LIGHTNESS), (hi + lo + 1) // 2)
    r, g, b = (f[..., c].astype(np.float64) for c in range(3))
    np.testing.assert_array_equal(_codes(orc, img, itype, GRAY_LUMINANCE), _round(0.3 * r + 0.59 * g + 0.11 * b))
    np.testing.assert_array_equal(_codes(orc, img, itype, GRAY_MIXER, (1.0, 2.0, 1.0)),
                                  _round(0.25 * r + 0.5 * g + 0.25 * b))


def test_lstar_known_answers(orc):
    # neutral grays: Y = v / 255 (the sRGB luminance row sums to 1)
    v = np.array([0, 1, 2, 46, 118, 128, 200, 255], np.uint8)
    img = np.repeat(np.repeat(v[None, :, None], 3, axis=2), 1, axis=0)
    y = v / 255.0
    L = np.where(y < 216 / 24389, 24389 / 27 * y, 116 * np.cbrt(y) - 16)
    np.testing.assert_array_equal(_codes(orc, img, U8, GRAY_LSTAR)[0], _round(255 * (L / 100)))
    assert _codes(orc, img, U8, GRAY_LSTAR)[0, -1] == 255 and _codes(orc, img, U8, GRAY_LSTAR)[0, 0] == 0
    # L* = 50 at Y = 0.1842: the 8-bit gray 47 (Y = 0.1843) projects to 128
    img47 = np.full((1, 1, 3), 47, np.uint8)
    assert _codes(orc, img47, U8, GRAY_LSTAR)[0, 0] == 128
    # primed: the gamma 1/0.45 first
    lin = y ** (1 / 0.45)
    Lp = np.where(lin < 216 / 24389, 24389 / 27 * lin, 116 * np.cbrt(lin) - 16)
    np.testing.assert_array_equal(_codes(orc, img, U8, GRAY_PLSTAR)[0], _round(255 * (Lp / 100)))


def test_mixer_rejects_bad_weights(orc):
    img = np.zeros((4, 4, 3), np.uint8)
    with pytest.raises(ValueError):
        _codes(orc, img, U8, GRAY_MIXER, (0.0, 0.0, 0.0))
    with pytest.raises(ValueError):
        _codes(orc, img, U8, GRAY_MIXER, (1.0, -1.0, 1.0))
