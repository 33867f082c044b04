"""Writes tests/golden/colorspace_lcms.npz: the colour transforms enfuse and
enblend build for images without an ICC profile (src/enfuse.cc:2385-2498:
cmsCreate_sRGBProfile, cmsCreateXYZProfile, cmsCreateLab2Profile at the sRGB
profile's media white point, INTENT_PERCEPTUAL, cmsFLAGS_NOCACHE |
cmsFLAGS_HIGHRESPRECALC), evaluated by the system's LittleCMS 2 runtime
(liblcms2.so.2) through ctypes on a grid of RGB values: RGB -> XYZ, RGB -> Lab,
and back Lab -> RGB, XYZ -> RGB; and for 16-bit codes the scaled Lab values
the pyramid holds and the RGB LittleCMS gives back for them.  Run in a container that has the library;
the fixture is data (inputs and LittleCMS's outputs)."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _lib():
    L = ctypes.CDLL("liblcms2.so.2")
    vp = ctypes.c_void_p
    for n in ("cmsCreate_sRGBProfile", "cmsCreateXYZProfile", "cmsCreateLab2Profile", "cmsCreateTransform",
              "cmsReadTag"):
        getattr(L, n).restype = vp
    L.cmsCreateLab2Profile.argtypes = [vp]
    L.cmsCreateTransform.argtypes = [vp, ctypes.c_uint32, vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
    L.cmsDoTransform.argtypes = [vp, vp, vp, ctypes.c_uint32]
    L.cmsReadTag.argtypes = [vp, ctypes.c_uint32]
    L.cmsXYZ2xyY.argtypes = [vp, vp]
    return L


def _fmt(colorspace):
    # TYPE_*_DBL: FLOAT_SH(1) | COLORSPACE_SH(cs) | CHANNELS_SH(3) | BYTES_SH(0)
    return (1 << 22) | (colorspace << 16) | (3 << 3)


PT_RGB, PT_XYZ, PT_Lab = 4, 9, 10
INTENT_PERCEPTUAL = 0
FLAGS = 0x0040 | 0x0400          # cmsFLAGS_NOCACHE | cmsFLAGS_HIGHRESPRECALC
SIG_WTPT = 0x77747074            # cmsSigMediaWhitePointTag


def main():
    L = _lib()
    srgb = L.cmsCreate_sRGBProfile()
    xyzp = L.cmsCreateXYZProfile()
    white = (ctypes.c_double * 3)()
    L.cmsXYZ2xyY(ctypes.cast(white, ctypes.c_void_p), L.cmsReadTag(srgb, SIG_WTPT))
    labp = L.cmsCreateLab2Profile(ctypes.cast(white, ctypes.c_void_p))
    t_xyz = L.cmsCreateTransform(srgb, _fmt(PT_RGB), xyzp, _fmt(PT_XYZ), INTENT_PERCEPTUAL, FLAGS)
    t_xyz_inv = L.cmsCreateTransform(xyzp, _fmt(PT_XYZ), srgb, _fmt(PT_RGB), INTENT_PERCEPTUAL, FLAGS)
    t_lab = L.cmsCreateTransform(srgb, _fmt(PT_RGB), labp, _fmt(PT_Lab), INTENT_PERCEPTUAL, FLAGS)
    t_lab_inv = L.cmsCreateTransform(labp, _fmt(PT_Lab), srgb, _fmt(PT_RGB), INTENT_PERCEPTUAL, FLAGS)

    def run(t, x):
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros_like(x)
        L.cmsDoTransform(t, x.ctypes.data, out.ctypes.data, len(x))
        return out

    g = np.linspace(0.0, 1.0, 9)
    grid = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    rng = np.random.default_rng(2024)
    special = np.array([[1, 1, 1], [0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0.04045] * 3, [0.04] * 3,
                        [0.0031308] * 3, [0.5] * 3], np.float64)
    rgb = np.concatenate([special, grid, rng.random((300, 3))])
    xyz = run(t_xyz, rgb)
    lab = run(t_lab, rgb)
    rgb_from_lab = run(t_lab_inv, lab)
    rgb_from_xyz = run(t_xyz_inv, xyz)
    # 16-bit codes: LittleCMS's Lab of code / 65535, the PyramidScale'd
    # values of it (s = 32768, src/fixmath.h:486-525), and LittleCMS's RGB of
    # the Lab those scaled values stand for
    codes = np.concatenate([np.array([[65535] * 3, [0] * 3, [65535, 0, 0], [0, 65535, 0], [0, 0, 65535],
                                      [32768] * 3, [2651] * 3, [1] * 3], np.int64),
                            rng.integers(0, 65536, size=(600, 3))]).astype(np.uint16)
    lab16 = run(t_lab, codes / 65535.0)
    s = 32768.0
    q16 = np.stack([np.round(s * lab16[:, 0] / 100.0), np.round(s * (200.0 + lab16[:, 1]) / 400.0),
                    np.round(s * (200.0 + lab16[:, 2]) / 400.0)], axis=1).astype(np.uint16)
    lab_q = np.stack([q16[:, 0] * 100.0 / s, q16[:, 1] * 400.0 / s - 200.0, q16[:, 2] * 400.0 / s - 200.0], axis=1)
    rgb_from_q16 = run(t_lab_inv, lab_q)
    np.savez_compressed(os.path.join(HERE, "colorspace_lcms.npz"), rgb=rgb, xyz=xyz, lab=lab,
                        rgb_from_lab=rgb_from_lab, rgb_from_xyz=rgb_from_xyz, codes16=codes, lab16=lab16,
                        q16=q16, rgb_from_q16=rgb_from_q16,
                        lab_white_xyY=np.array(list(white)), lcms_version=np.array(L.cmsGetEncodedCMMversion()))
    print("wrote", len(rgb), "samples; Lab white xyY", list(white))


if __name__ == "__main__":
    main()
