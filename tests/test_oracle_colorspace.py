"""Pins the oracle's blend colour spaces (oracle/oracle_colorspace.cpp,
src/fixmath.h:429-1330) against LittleCMS 2's own transforms for the sRGB
input profile (tests/golden/colorspace_lcms.npz, written by
tests/golden/make_colorspace_fixtures.py) and against closed forms.
LittleCMS evaluates its pipelines in single precision, so agreement is to
its float rounding; Luv is VIGRA's functor on LittleCMS's XYZ, pinned here
through its lightness (the Lab L*) and its chromaticities u', v' of the
XYZ LittleCMS gives (VIGRA's own bits: parity unpinned)."""
import os

import numpy as np
import pytest

from oracle_binding import CS_CIELAB, CS_CIELUV, U8, U16, load_oracle

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "colorspace_lcms.npz"))
D50 = np.array([0.9642, 1.0, 0.8249])


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


def test_lab_forward_matches_lcms(orc):
    got = orc.colorspace_forward(CS_CIELAB, FIX["rgb"])
    err = np.abs(got - FIX["lab"]).max()
    assert err < 5e-5, err


def test_lab_inverse_matches_lcms(orc):
    got = orc.colorspace_inverse(CS_CIELAB, FIX["lab"])
    err = np.abs(got - FIX["rgb_from_lab"]).max()
    assert err < 5e-6, err


def test_matrix_and_transfer_curve_match_lcms_xyz(orc):
    M = orc.rgb_to_xyz_matrix()
    rgb = FIX["rgb"]
    lin = np.where(rgb >= 0.04045, ((rgb + 0.055) / 1.055) ** 2.4, rgb / 12.92)
    err = np.abs(lin @ M.T - FIX["xyz"]).max()
    assert err < 5e-7, err
    # closed forms: RGB = 1 is the D50 white, the Y row sums to 1
    np.testing.assert_allclose(M.sum(axis=1), D50, rtol=0, atol=1e-12)
    np.testing.assert_allclose(M[1].sum(), 1.0, rtol=0, atol=1e-15)
    # the Lab profile's white is the sRGB profile's media white, D50
    xyY = FIX["lab_white_xyY"]
    np.testing.assert_allclose([xyY[0] * xyY[2] / xyY[1], xyY[2], (1 - xyY[0] - xyY[1]) * xyY[2] / xyY[1]], D50,
                               atol=1e-6)


def test_luv_lightness_and_chromaticity(orc):
    luv = orc.colorspace_forward(CS_CIELUV, FIX["rgb"])
    lab = FIX["lab"]
    # L* is the same function of Y in both spaces
    assert np.abs(luv[:, 0] - lab[:, 0]).max() < 5e-5
    # u* / 13 L* + u'n = u' of the (LittleCMS) XYZ, likewise v'
    xyz = FIX["xyz"]
    den = xyz[:, 0] + 15 * xyz[:, 1] + 3 * xyz[:, 2]
    ok = luv[:, 0] > 1.0
    up = 4 * xyz[ok, 0] / den[ok]
    vp = 9 * xyz[ok, 1] / den[ok]
    assert np.abs(luv[ok, 1] / (13 * luv[ok, 0]) + 0.197839 - up).max() < 1e-6
    assert np.abs(luv[ok, 2] / (13 * luv[ok, 0]) + 0.468342 - vp).max() < 1e-6
    # VIGRA's white chromaticities are those of its D65 white
    # (0.950456, 1, 1.088754)
    den_n = 0.950456 + 15 + 3 * 1.088754
    assert abs(4 * 0.950456 / den_n - 0.197839) < 1e-6
    assert abs(9 / den_n - 0.468342) < 1e-6
    # black is (0, 0, 0) (VIGRA's Y == 0 branch)
    assert np.all(luv[1] == 0.0)


@pytest.mark.parametrize("space", [CS_CIELAB, CS_CIELUV])
def test_round_trip_and_known_answers(orc, space):
    rgb = FIX["rgb"]
    back = orc.colorspace_inverse(space, orc.colorspace_forward(space, rgb))
    # the two pieces of the sRGB curve meet at 0.04045 with a gap of ~1e-8:
    # values in it come back off by that much
    np.testing.assert_allclose(back, rgb, rtol=0, atol=1e-7)
    white = orc.colorspace_forward(space, [[1.0, 1.0, 1.0]])[0]
    assert abs(white[0] - 100.0) < 1e-12
    if space == CS_CIELAB:
        assert np.abs(white[1:]).max() < 1e-12
        gray = orc.colorspace_forward(space, np.repeat(np.linspace(0, 1, 11)[:, None], 3, axis=1))
        assert np.abs(gray[:, 1:]).max() < 1e-12
        # 18% gray reflectance has L* = 116 * 0.18^(1/3) - 16
        y18 = 0.18
        c = ((y18 ** (1 / 2.4)) * 1.055 - 0.055)
        assert abs(orc.colorspace_forward(space, [[c, c, c]])[0][0] - (116 * y18 ** (1 / 3) - 16)) < 1e-9


@pytest.mark.parametrize("itype", [U8, U16])
@pytest.mark.parametrize("space", [CS_CIELAB, CS_CIELUV])
def test_scaled_images(orc, itype, space):
    """PyramidScale (src/fixmath.h:486-525): L' = round(s L / 100),
    a' = round(s (200 + a) / 400) with s = 128 (8-bit) or 32768 (16-bit);
    against the same scaling of LittleCMS's Lab (or of the Luv of its XYZ),
    equal up to one unit where the two sit within LittleCMS's float rounding
    of a half; and back through the inverse scaling within the quantisation
    of the scaled values."""
    mx, s, dt = (255, 128.0, np.uint8) if itype == U8 else (65535, 32768.0, np.uint16)
    rng = np.random.default_rng(itype + space)
    img = rng.integers(0, mx + 1, size=(37, 41, 3)).astype(dt)
    img[0, 0] = mx
    img[0, 1] = 0
    got = orc.rgb_to_blend_space(space, img, itype).astype(np.int64)
    assert tuple(got[0, 0]) == ((int(s), int(s) // 2, int(s) // 2) if space == CS_CIELAB else tuple(got[0, 0]))
    assert tuple(got[0, 1]) == (0, int(s) // 2, int(s) // 2)
    v = orc.colorspace_forward(space, img.reshape(-1, 3) / mx)
    want = np.stack([np.round(s * v[:, 0] / 100.0), np.round(s * (200.0 + v[:, 1]) / 400.0),
                     np.round(s * (200.0 + v[:, 2]) / 400.0)], axis=1).astype(np.int64)
    np.testing.assert_array_equal(got.reshape(-1, 3), want)
    mask = np.where(rng.random(img.shape[:2]) < 0.7, 255, 0).astype(np.uint8)
    back = orc.blend_space_to_rgb_if(space, got.astype(dt), mask, np.zeros_like(img), itype)
    assert np.all(back[mask == 0] == 0)
    sel = mask != 0
    # inverse scaling L = q 100 / s, a = q 400 / s - 200, then the RGB value
    # times the type's maximum, rounded half away from zero and clamped
    q = got[sel].astype(np.float64)
    v = np.stack([q[:, 0] * 100.0 / s, q[:, 1] * 400.0 / s - 200.0, q[:, 2] * 400.0 / s - 200.0], axis=1)
    r = mx * orc.colorspace_inverse(space, v)
    want_back = np.clip(np.where(r >= 0, np.floor(r + 0.5), np.ceil(r - 0.5)), 0, mx)
    np.testing.assert_array_equal(back[sel], want_back)
    if itype == U16:
        # a scaled unit is 1/327.68 of L* and 1/81.92 of a*, b*: half of one
        # moves linear light by well under 3e-4 (dark, saturated colours
        # move most in sRGB codes)
        def lin(x):
            return np.where(x >= 0.04045, ((x + 0.055) / 1.055) ** 2.4, x / 12.92)

        d = np.abs(lin(back[sel] / mx) - lin(img[sel] / mx))
        assert d.max() < 3e-4, d.max()


def test_scaled_values_against_lcms_lab(orc):
    """The 16-bit scaling of the oracle's Lab and of LittleCMS's Lab of the
    same RGB values agree within one unit (LittleCMS's float rounding moves
    values that sit at a half)."""
    s = 32768.0
    rgb = FIX["rgb"]
    mine = orc.colorspace_forward(CS_CIELAB, rgb)

    def scale(v):
        return np.stack([np.round(s * v[:, 0] / 100.0), np.round(s * (200.0 + v[:, 1]) / 400.0),
                         np.round(s * (200.0 + v[:, 2]) / 400.0)], axis=1)

    d = np.abs(scale(mine) - scale(FIX["lab"]))
    assert d.max() <= 1
    assert np.mean(d == 0) > 0.99


def _from_real_u16(r):
    r = 65535.0 * r
    return np.clip(np.where(r >= 0, np.floor(r + 0.5), np.ceil(r - 0.5)), 0, 65535).astype(np.int64)


def test_16bit_codes_against_lcms(orc):
    """16-bit codes through PyramidScale'd Lab and back, against
    LittleCMS's Lab of the same codes and its RGB of the scaled values:
    equal but for one unit where LittleCMS's float rounding crosses a half."""
    codes = FIX["codes16"].reshape(1, -1, 3)
    q = orc.rgb_to_blend_space(CS_CIELAB, codes, U16).reshape(-1, 3).astype(np.int64)
    d = np.abs(q - FIX["q16"].astype(np.int64))
    assert d.max() <= 1 and np.mean(d == 0) > 0.995
    mask = np.full(codes.shape[:2], 255, np.uint8)
    back = orc.blend_space_to_rgb_if(CS_CIELAB, FIX["q16"].reshape(1, -1, 3), mask, np.zeros_like(codes), U16)
    d = np.abs(back.reshape(-1, 3).astype(np.int64) - _from_real_u16(FIX["rgb_from_q16"]))
    assert d.max() <= 1 and np.mean(d == 0) > 0.995


# ---- CIECAM02 (JCh blend space, src/fixmath.h:1336-1990) ---------------------

CAM = np.load(os.path.join(HERE, "golden", "ciecam_lcms.npz"))


def test_ciecam_viewing_conditions_are_the_references():
    """The fixture was written with the reference's viewing conditions
    (src/enblend.cc:2235-2244): white 100 x D50."""
    np.testing.assert_allclose(CAM["white"], 100.0 * D50, rtol=1e-4)


def test_ciecam_forward_matches_lcms(orc):
    """LittleCMS's own cmsCIECAM02Forward on the same XYZ: J to 1e-12,
    C and h to 1e-9 relative (where C is not tiny)."""
    got = orc.ciecam_forward(CAM["xyz"])
    ref = CAM["jch"]
    assert np.abs(got[:, 0] - ref[:, 0]).max() < 1e-12
    m = ref[:, 1] > 1.0
    assert (np.abs(got[m, 1:] - ref[m, 1:]) / ref[m, 1:]).max() < 1e-9
    assert np.abs(got[:, 1] - ref[:, 1]).max() < 1e-7


def test_ciecam_reverse_matches_lcms(orc):
    """cmsCIECAM02Reverse over the forward results and over JCh all round the
    hue circle: equal to 1e-12 relative.  (The inverse HPE matrix entry that
    LittleCMS evaluates with the opposite sign of the published one, and its
    pi = 3.141592654, are what brings this to rounding level: with the
    published sign alone the XYZ differ by about 2e-5.)"""
    got = orc.ciecam_reverse(CAM["jch_in"])
    ref = CAM["xyz_back"]
    assert (np.abs(got - ref) / np.abs(ref).max(axis=1, keepdims=True)).max() < 1e-12


def test_ciecam_known_answers(orc):
    """The adopted white has J = 100 exactly (A = Aw); black has J = C = 0 up
    to rounding; hue angles lie in [0, 360)."""
    jw = orc.ciecam_forward(100.0 * D50[None, :])[0]
    assert jw[0] == 100.0 and jw[1] < 0.01
    jb = orc.ciecam_forward(np.zeros((1, 3)))[0]
    assert abs(jb[0]) < 1e-20 and abs(jb[1]) < 1e-6
    j = orc.ciecam_forward(CAM["xyz"])
    assert (j[:, 2] >= 0).all() and (j[:, 2] < 360).all()


@pytest.mark.parametrize("itype", [U8, U16])
def test_jch_level0_scale_and_round_trip(orc, itype):
    """PyramidScaleJCh: the J plane of white holds 100 x 2^(IntegerBits-8) x
    2^7 exactly, black is all zeros; forward then inverse under a full mask
    gives the codes back within the asymmetry of LittleCMS's forward and
    reverse models (its reverse is not the exact inverse of its forward)."""
    m = 255 if itype == U8 else 65535
    rng = np.random.default_rng(itype)
    img = rng.integers(0, m + 1, size=(60, 80, 3)).astype(np.uint8 if itype == U8 else np.uint16)
    img[0, 0] = m
    img[0, 1] = 0
    lv = orc.rgb_to_jch_level0(img, itype)
    s = 2 if itype == U8 else 512
    assert lv[0, 0, 0] == 100 * s * 128
    assert (lv[:, 0, 1] == 0).all()
    back = orc.jch_level0_to_rgb_if(lv, np.full(img.shape[:2], 255, np.uint8), np.zeros_like(img), itype)
    d = np.abs(back.astype(np.int64) - img.astype(np.int64))
    assert d.max() <= (1 if itype == U8 else 32), d.max()
    assert np.median(d) <= (0 if itype == U8 else 8)
