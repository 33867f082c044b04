"""Writes tests/golden/ciecam_lcms.npz: the CIECAM02 model as the reference
sets it up for its JCh blend space (src/enblend.cc:2235-2244,
src/enfuse.cc's identical block: white point 100 x D50, Yb 20, La 31.83,
average surround, D 1.0), evaluated by the system's LittleCMS 2 runtime
(liblcms2.so.2: cmsCIECAM02Init / cmsCIECAM02Forward / cmsCIECAM02Reverse)
through ctypes: JCh of XYZ samples (100 x LittleCMS's sRGB -> XYZ of an RGB
grid, as rgb_to_jch feeds it, src/fixmath.h:1443-1458) and XYZ of JCh
samples.  Run in a container that has the library; the fixture is data
(inputs and LittleCMS's outputs)."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class VC(ctypes.Structure):
    """cmsViewingConditions."""
    _fields_ = [("X", ctypes.c_double), ("Y", ctypes.c_double), ("Z", ctypes.c_double), ("Yb", ctypes.c_double),
                ("La", ctypes.c_double), ("surround", ctypes.c_uint32), ("D_value", ctypes.c_double)]


def _fmt(colorspace):
    return (1 << 22) | (colorspace << 16) | (3 << 3)   # TYPE_*_DBL


def main():
    L = ctypes.CDLL("liblcms2.so.2")
    vp = ctypes.c_void_p
    L.cmsD50_XYZ.restype = ctypes.POINTER(ctypes.c_double * 3)
    L.cmsCIECAM02Init.restype = vp
    L.cmsCIECAM02Init.argtypes = [vp, ctypes.POINTER(VC)]
    L.cmsCIECAM02Forward.argtypes = [vp, vp, vp]
    L.cmsCIECAM02Reverse.argtypes = [vp, vp, vp]
    for n in ("cmsCreate_sRGBProfile", "cmsCreateXYZProfile", "cmsCreateTransform"):
        getattr(L, n).restype = vp
    L.cmsCreateTransform.argtypes = [vp, ctypes.c_uint32, vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
    L.cmsDoTransform.argtypes = [vp, vp, vp, ctypes.c_uint32]

    d50 = list(L.cmsD50_XYZ().contents)
    vc = VC(100.0 * d50[0], 100.0 * d50[1], 100.0 * d50[2], 20.0, 31.83, 1, 1.0)   # AVG_SURROUND = 1
    cam = L.cmsCIECAM02Init(None, ctypes.byref(vc))
    assert cam

    t_xyz = L.cmsCreateTransform(L.cmsCreate_sRGBProfile(), _fmt(4), L.cmsCreateXYZProfile(), _fmt(9), 0,
                                 0x0040 | 0x0400)
    g = np.linspace(0.0, 1.0, 9)
    grid = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    rng = np.random.default_rng(2025)
    special = np.array([[1, 1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0.5] * 3, [0.01] * 3, [1e-4, 0, 0]], np.float64)
    rgb = np.concatenate([special, grid[1:], rng.random((400, 3))])   # grid[0] is black (J = 0)
    xyz01 = np.zeros_like(rgb)
    L.cmsDoTransform(t_xyz, rgb.ctypes.data, xyz01.ctypes.data, len(rgb))
    xyz = 100.0 * xyz01

    def forward(x):
        out = np.zeros_like(x)
        for i in range(len(x)):
            xi = np.ascontiguousarray(x[i])
            oi = np.zeros(3)
            L.cmsCIECAM02Forward(cam, xi.ctypes.data, oi.ctypes.data)
            out[i] = oi
        return out

    def reverse(j):
        out = np.zeros_like(j)
        for i in range(len(j)):
            ji = np.ascontiguousarray(j[i])
            oi = np.zeros(3)
            L.cmsCIECAM02Reverse(cam, ji.ctypes.data, oi.ctypes.data)
            out[i] = oi
        return out

    jch = forward(xyz)
    # JCh samples over the blend space's range, hue all round
    jch_in = np.concatenate([jch, np.stack([rng.uniform(1.0, 100.0, 300), rng.uniform(0.0, 60.0, 300),
                                            rng.uniform(0.0, 360.0, 300)], axis=1)])
    xyz_back = reverse(jch_in)
    np.savez_compressed(os.path.join(HERE, "ciecam_lcms.npz"), rgb=rgb, xyz=xyz, jch=jch, jch_in=jch_in,
                        xyz_back=xyz_back, white=np.array([vc.X, vc.Y, vc.Z]),
                        lcms_version=np.array(L.cmsGetEncodedCMMversion()))
    print("wrote", len(xyz), "forward and", len(jch_in), "reverse samples")


if __name__ == "__main__":
    main()
