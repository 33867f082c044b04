"""ctypes binding of the oracle (oracle/liboracle.so): the CPU restatement
used as the checker.  Builds it with oracle/Makefile when missing."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

I16, I32, F64 = 0, 1, 2
U8, U16, F32, U32 = 10, 11, 12, 13
NP_P = {I16: np.int16, I32: np.int32, F64: np.float64}
NP_I = {U8: np.uint8, U16: np.uint16, F32: np.float32, U32: np.uint32}

_vp, _i, _l = ctypes.c_void_p, ctypes.c_int, ctypes.c_long
_SIGS = {
    "orc_reduce": [_i, _i, _vp, _i, _i, _l, _vp, _l, _vp, _i, _i, _l, _vp, _l],
    "orc_expand": [_i, _i, _i, _vp, _i, _i, _l, _vp, _i, _i, _l],
    "orc_auto_levels": [_i, _i],
    "orc_to_pyramid": [_i, _i, _i, _i, _vp, _i, _i, _l, _vp, _l, _l],
    "orc_from_pyramid_if": [_i, _i, _i, _i, _vp, _i, _i, _l, _l, _vp, _l, _vp, _l, _i, ctypes.c_uint64],
    "orc_dither_rand": [ctypes.c_uint64, _i, _i, _i, _i],
    "orc_blend_level": [_i, _i, _i, _i, _i, _vp, _l, _vp, _l, _l, _vp, _l, _l, ctypes.c_double],
    "orc_enblend_pair": [_i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, ctypes.c_uint64],
    "orc_pyramid": [_i, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, ctypes.POINTER(_l)],
    "orc_collapse": [_i, _i, _i, _i, _i, _i, _vp],
    "orc_num_threads": [],
    "orc_enfuse_mask": [_i, _vp, _vp, _i, _i, _vp, _vp],
    "orc_enfuse": [_i, _i, _i, _i, _vp, _vp, _vp, _i, _i, _vp, _vp, _i, ctypes.c_uint64],
    "orc_diff_forward": [_i, _i, _i, _vp, _vp, _i, _i, _vp],
    "orc_diff_backward": [_i, _i, _i, _vp, _vp, _i, _i, _vp, _vp, _vp],
    "orc_diff_expand_T": [_vp, _i, _i, _vp, _i, _i],
    "orc_diff_reduce_T": [_vp, _i, _i, _vp, _i, _i],
    "orc_nft_mask": [_vp, _l, _vp, _l, _i, _i, _i, _vp, _l],
    "orc_distance_transform": [_vp, _l, _i, _i, _i, _vp, _l],
    "orc_euclid_1d": [_vp, _i, _i, _vp],
    "orc_set_blend_colorspace": [_i],
    "orc_rgb_to_blend_space": [_i, _i, _i, _i, _vp, _vp],
    "orc_blend_space_to_rgb_if": [_i, _i, _i, _i, _vp, _vp, _vp],
    "orc_colorspace_forward": [_i, _vp, _l, _vp],
    "orc_ciecam_forward": [_vp, _l, _vp],
    "orc_ciecam_reverse": [_vp, _l, _vp],
    "orc_rgb_to_jch_level0": [_i, _i, _i, _vp, _vp, _l],
    "orc_jch_level0_to_rgb_if": [_i, _i, _i, _vp, _l, _vp, _vp],
    "orc_colorspace_inverse": [_i, _vp, _l, _vp],
    "orc_rgb_to_xyz_matrix": [_vp],
    "orc_local_stddev_f32": [_vp, _vp, _i, _i, _i, _i, _vp],
    "orc_nft_mask_periodic": [_vp, _l, _vp, _l, _i, _i, _i, _i, _vp, _l],
    "orc_assemble": [_i, _i, _i, _i, _i, _vp, _i, _vp, _vp, _vp, _vp],
    "orc_inspect_overlap": [_vp, _l, _vp, _l, _i, _i],
    "orc_mask_bounds": [_vp, _i, _i, _vp],
    "orc_enblend_step": [_i, _i, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, ctypes.c_uint64,
                         _vp, _vp],
    "orc_checkpoint": [_i, _i, _i, _i, _vp, _vp, _i, _vp, _vp],
    "orc_vigra_kernel": [ctypes.c_double, _i, _vp, _i],
}


class Layer(ctypes.Structure):
    """orc_layer / enbl_layer (same layout): one decoded input image."""
    _fields_ = [("data", ctypes.c_void_p), ("pitch", ctypes.c_long), ("w", ctypes.c_int), ("h", ctypes.c_int),
                ("x", ctypes.c_int), ("y", ctypes.c_int), ("ftype", ctypes.c_int), ("bands", ctypes.c_int)]


def layer_table(layers, ptr=lambda a: a.ctypes.data):
    """ctypes array of Layer for (samples (h, w, bands), x, y, ftype) tuples;
    `ptr` gives the address of the samples (host arrays by default)."""
    tab = (Layer * max(1, len(layers)))()
    for i, (a, x, y, ft) in enumerate(layers):
        h, w, bands = a.shape
        tab[i] = Layer(ptr(a), w * bands, w, h, x, y, ft, bands)
    return tab

EXP_GAUSSIAN, EXP_LORENTZIAN, EXP_HALFSINE, EXP_FULLSINE, EXP_BISQUARE = range(5)
(GRAY_AVERAGE, GRAY_LSTAR, GRAY_PLSTAR, GRAY_LIGHTNESS, GRAY_VALUE, GRAY_ANTI_VALUE, GRAY_LUMINANCE,
 GRAY_MIXER) = range(8)


CS_IDENTITY, CS_CIELAB, CS_CIELUV, CS_CIECAM = 0, 1, 2, 3


class EnfuseParams(ctypes.Structure):
    """orc_enfuse_params / enbl_enfuse_params (same layout)."""
    _fields_ = [("w_exposure", ctypes.c_double), ("w_saturation", ctypes.c_double),
                ("w_contrast", ctypes.c_double), ("w_entropy", ctypes.c_double),
                ("exposure_optimum", ctypes.c_double), ("exposure_width", ctypes.c_double),
                ("exposure_function", ctypes.c_int), ("contrast_window", ctypes.c_int),
                ("entropy_window", ctypes.c_int), ("hard_mask", ctypes.c_int),
                ("exposure_table", ctypes.c_void_p), ("edge_scale", ctypes.c_double),
                ("lce_scale", ctypes.c_double), ("lce_factor", ctypes.c_double),
                ("min_curvature", ctypes.c_double), ("exposure_cutoff", ctypes.c_int),
                ("entropy_cutoff", ctypes.c_int), ("exposure_lower_cutoff", ctypes.c_double),
                ("exposure_upper_cutoff", ctypes.c_double), ("entropy_lower_cutoff", ctypes.c_double),
                ("entropy_upper_cutoff", ctypes.c_double), ("gray_projector", ctypes.c_int),
                ("mixer_red", ctypes.c_double), ("mixer_green", ctypes.c_double), ("mixer_blue", ctypes.c_double)]


def enfuse_params(w_exposure=1.0, w_saturation=0.0, w_contrast=0.0, w_entropy=0.0,
                  exposure_optimum=0.5, exposure_width=0.2, exposure_function=EXP_GAUSSIAN,
                  contrast_window=5, entropy_window=3, exposure_table=None, hard_mask=False,
                  edge_scale=0.0, lce_scale=0.0, lce_factor=0.0, min_curvature=0.0,
                  exposure_cutoff=None, entropy_cutoff=None, gray_projector=0, mixer=(0.0, 0.0, 0.0)):
    p = EnfuseParams(w_exposure, w_saturation, w_contrast, w_entropy, exposure_optimum, exposure_width,
                     exposure_function, contrast_window, entropy_window, int(hard_mask), None,
                     edge_scale, lce_scale, lce_factor, min_curvature)
    p.gray_projector = gray_projector
    p.mixer_red, p.mixer_green, p.mixer_blue = mixer
    if exposure_cutoff is not None:   # (lower, upper), instantiated for the scalar type
        p.exposure_cutoff = 1
        p.exposure_lower_cutoff, p.exposure_upper_cutoff = exposure_cutoff
    if entropy_cutoff is not None:
        p.entropy_cutoff = 1
        p.entropy_lower_cutoff, p.entropy_upper_cutoff = entropy_cutoff
    if exposure_table is not None:
        tab = np.ascontiguousarray(exposure_table, np.float64)
        p.exposure_table = tab.ctypes.data
        p._keep = tab
    return p


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


class Oracle:
    def __init__(self, lib):
        self.lib = lib

    def reduce(self, src, alpha, wrap, ptype):
        """src: (h, w) array of the pyramid type; alpha: (h, w) uint8 or None."""
        sh, sw = src.shape
        dh, dw = (sh + 1) // 2, (sw + 1) // 2
        src = np.ascontiguousarray(src, NP_P[ptype])
        dst = np.zeros((dh, dw), NP_P[ptype])
        da = np.zeros((dh, dw), np.uint8) if alpha is not None else None
        if alpha is not None:
            alpha = np.ascontiguousarray(alpha, np.uint8)
        rc = self.lib.orc_reduce(ptype, int(wrap), _p(src), sw, sh, sw, _p(alpha), sw,
                                 _p(dst), dw, dh, dw, _p(da), dw)
        if rc:
            raise ValueError(f"orc_reduce rc={rc}")
        return (dst, da) if alpha is not None else dst

    def expand(self, src, dst, add, wrap, ptype):
        """Returns a new array: dst combined with expand(src)."""
        sh, sw = src.shape
        dh, dw = dst.shape
        src = np.ascontiguousarray(src, NP_P[ptype])
        out = np.array(dst, dtype=NP_P[ptype], copy=True)
        rc = self.lib.orc_expand(ptype, int(add), int(wrap), _p(src), sw, sh, sw, _p(out), dw, dh, dw)
        if rc:
            raise ValueError(f"orc_expand rc={rc}")
        return out

    def auto_levels(self, w, h):
        return self.lib.orc_auto_levels(w, h)

    def pyramid(self, img, alpha, levels, wrap, laplacian, itype):
        """img: (h, w, C) interleaved; returns a list of (C, h_l, w_l) levels."""
        h, w, C = img.shape
        ptype = {U8: I16, U16: I32, F32: F64}[itype]
        dims = []
        tw, th = w, h
        total = 0
        for _ in range(levels):
            dims.append((th, tw))
            total += C * tw * th
            tw, th = (tw + 1) // 2, (th + 1) // 2
        out = np.zeros(total, NP_P[ptype])
        n = _l()
        img = np.ascontiguousarray(img, NP_I[itype])
        a = None if alpha is None else np.ascontiguousarray(alpha, np.uint8)
        rc = self.lib.orc_pyramid(int(laplacian), itype, C, w, h, levels, int(wrap), _p(img), _p(a), _p(out),
                                  ctypes.byref(n))
        if rc:
            raise ValueError(f"orc_pyramid rc={rc}")
        res, off = [], 0
        for (lh, lw) in dims:
            res.append(out[off: off + C * lh * lw].reshape(C, lh, lw).copy())
            off += C * lh * lw
        return res

    def collapse(self, levels_list, wrap, ptype):
        C = levels_list[0].shape[0]
        h, w = levels_list[0].shape[1:]
        flat = np.concatenate([np.ascontiguousarray(l, NP_P[ptype]).ravel() for l in levels_list])
        rc = self.lib.orc_collapse(ptype, C, w, h, len(levels_list), int(wrap), _p(flat))
        if rc:
            raise ValueError(f"orc_collapse rc={rc}")
        res, off = [], 0
        for l in levels_list:
            n = l.size
            res.append(flat[off: off + n].reshape(l.shape).copy())
            off += n
        return res

    def to_pyramid(self, img, itype, ptype, frac):
        h, w, C = img.shape
        out = np.zeros((C, h, w), NP_P[ptype])
        img = np.ascontiguousarray(img, NP_I[itype])
        rc = self.lib.orc_to_pyramid(itype, ptype, frac, C, _p(img), w, h, w * C, _p(out), w, w * h)
        if rc:
            raise ValueError(f"orc_to_pyramid rc={rc}")
        return out

    def from_pyramid_if(self, planes, mask, img, ptype, itype, frac, dither, seed):
        C, h, w = planes.shape
        out = np.array(img, dtype=NP_I[itype], copy=True)
        planes = np.ascontiguousarray(planes, NP_P[ptype])
        mask = np.ascontiguousarray(mask, np.uint8)
        rc = self.lib.orc_from_pyramid_if(ptype, itype, frac, C, _p(planes), w, h, w, w * h, _p(mask), w,
                                          _p(out), w * C, dither, seed)
        if rc:
            raise ValueError(f"orc_from_pyramid_if rc={rc}")
        return out

    def dither_rand(self, seed, x, y, c, frac):
        return self.lib.orc_dither_rand(seed, x, y, c, frac)

    def blend_level(self, mask, white, black, mtype, ptype, white_value):
        C, h, w = white.shape
        out = np.array(black, dtype=NP_P[ptype], copy=True)
        mask = np.ascontiguousarray(mask, NP_P[mtype])
        white = np.ascontiguousarray(white, NP_P[ptype])
        rc = self.lib.orc_blend_level(mtype, ptype, C, w, h, _p(mask), w, _p(white), w, w * h, _p(out), w, w * h,
                                      float(white_value))
        if rc:
            raise ValueError(f"orc_blend_level rc={rc}")
        return out

    def enblend_pair(self, white, walpha, black, balpha, mask, levels, wrap, itype, dither, seed):
        h, w, C = white.shape
        b = np.array(black, dtype=NP_I[itype], copy=True)
        ba = np.array(balpha, dtype=np.uint8, copy=True)
        white = np.ascontiguousarray(white, NP_I[itype])
        walpha = np.ascontiguousarray(walpha, np.uint8)
        mask = np.ascontiguousarray(mask, np.uint8)
        rc = self.lib.orc_enblend_pair(itype, C, w, h, levels, int(wrap), _p(white), _p(walpha), _p(b), _p(ba),
                                       _p(mask), dither, seed)
        if rc:
            raise ValueError(f"orc_enblend_pair rc={rc}")
        return b, ba


def _enfuse_mask(self, img, alpha, itype, prm):
    h, w, _ = img.shape
    out = np.zeros((h, w), np.float32)
    img = np.ascontiguousarray(img, NP_I[itype])
    alpha = np.ascontiguousarray(alpha, np.uint8)
    rc = self.lib.orc_enfuse_mask(itype, _p(img), _p(alpha), w, h, ctypes.byref(prm), _p(out))
    if rc:
        raise ValueError(f"orc_enfuse_mask rc={rc}")
    return out


def _enfuse(self, imgs, alphas, itype, prm, levels, wrap=False, dither=1, seed=0):
    h, w, _ = imgs[0].shape
    K = len(imgs)
    imgs = [np.ascontiguousarray(i, NP_I[itype]) for i in imgs]
    alphas = [np.ascontiguousarray(a, np.uint8) for a in alphas]
    ip = (ctypes.c_void_p * K)(*[i.ctypes.data for i in imgs])
    ap = (ctypes.c_void_p * K)(*[a.ctypes.data for a in alphas])
    out = np.zeros((h, w, 3), NP_I[itype])
    oa = np.zeros((h, w), np.uint8)
    rc = self.lib.orc_enfuse(itype, w, h, K, ip, ap, ctypes.byref(prm), levels, int(wrap), _p(out), _p(oa),
                             dither, seed)
    if rc:
        raise ValueError(f"orc_enfuse rc={rc}")
    return out, oa


Oracle.enfuse_mask = _enfuse_mask
Oracle.enfuse = _enfuse


def _diff_forward(self, imgs, prm, levels, n=5):
    """Relaxed differentiable enfuse: imgs (K, h, w, 3) float64, prm 5 values."""
    imgs = np.ascontiguousarray(imgs, np.float64)
    K, h, w, _ = imgs.shape
    prm = np.ascontiguousarray(prm, np.float64)
    out = np.zeros((h, w, 3), np.float64)
    rc = self.lib.orc_diff_forward(w, h, K, _p(imgs), _p(prm), levels, n, _p(out))
    if rc:
        raise ValueError(f"orc_diff_forward rc={rc}")
    return out


def _diff_backward(self, imgs, prm, levels, gout, n=5):
    imgs = np.ascontiguousarray(imgs, np.float64)
    K, h, w, _ = imgs.shape
    prm = np.ascontiguousarray(prm, np.float64)
    gout = np.ascontiguousarray(gout, np.float64)
    gimgs = np.zeros_like(imgs)
    gprm = np.zeros(5, np.float64)
    rc = self.lib.orc_diff_backward(w, h, K, _p(imgs), _p(prm), levels, n, _p(gout), _p(gimgs), _p(gprm))
    if rc:
        raise ValueError(f"orc_diff_backward rc={rc}")
    return gimgs, gprm


def _expand_T(self, g, cw, ch):
    g = np.ascontiguousarray(g, np.float64)
    fh, fw = g.shape
    gc = np.zeros((ch, cw), np.float64)
    self.lib.orc_diff_expand_T(_p(g), fw, fh, _p(gc), cw, ch)
    return gc


def _reduce_T(self, g, sw, sh):
    g = np.ascontiguousarray(g, np.float64)
    dh, dw = g.shape
    gf = np.zeros((sh, sw), np.float64)
    self.lib.orc_diff_reduce_T(_p(g), dw, dh, _p(gf), sw, sh)
    return gf


def _distance_transform(self, src, norm=2):
    """Float distance of every pixel to the nearest non-zero pixel."""
    src = np.ascontiguousarray(src, np.uint8)
    h, w = src.shape
    out = np.zeros((h, w), np.float32)
    rc = self.lib.orc_distance_transform(_p(src), w, w, h, norm, _p(out), w)
    if rc:
        raise ValueError(f"orc_distance_transform rc={rc}")
    return out


def _euclid_1d(self, f, planted=0):
    """One Euclidean lower-envelope transform of the float values f."""
    f = np.ascontiguousarray(f, np.float32)
    out = np.zeros(f.shape, np.float32)
    rc = self.lib.orc_euclid_1d(_p(f), f.size, planted, _p(out))
    if rc:
        raise ValueError(f"orc_euclid_1d rc={rc}")
    return out


def _local_stddev_f32(self, gray, mask, n, planted=0):
    """Float localStdDevIf; pixels it does not write stay 0."""
    gray = np.ascontiguousarray(gray, np.float32)
    mask = np.ascontiguousarray(mask, np.uint8)
    h, w = gray.shape
    out = np.zeros((h, w), np.float32)
    rc = self.lib.orc_local_stddev_f32(_p(gray), _p(mask), w, h, n, planted, _p(out))
    if rc:
        raise ValueError(f"orc_local_stddev_f32 rc={rc}")
    return out


def _set_blend_colorspace(self, space):
    if self.lib.orc_set_blend_colorspace(space):
        raise ValueError(f"bad colour space {space}")


def _rgb_to_blend_space(self, space, img, itype):
    img = np.ascontiguousarray(img, NP_I[itype])
    h, w, _ = img.shape
    out = np.zeros_like(img)
    if self.lib.orc_rgb_to_blend_space(space, itype, w, h, _p(img), _p(out)):
        raise ValueError("orc_rgb_to_blend_space")
    return out


def _blend_space_to_rgb_if(self, space, src, mask, img, itype):
    src = np.ascontiguousarray(src, NP_I[itype])
    mask = np.ascontiguousarray(mask, np.uint8)
    out = np.array(img, NP_I[itype], copy=True)
    h, w, _ = src.shape
    if self.lib.orc_blend_space_to_rgb_if(space, itype, w, h, _p(src), _p(mask), _p(out)):
        raise ValueError("orc_blend_space_to_rgb_if")
    return out


def _colorspace_forward(self, space, rgb):
    rgb = np.ascontiguousarray(rgb, np.float64).reshape(-1, 3)
    out = np.zeros_like(rgb)
    if self.lib.orc_colorspace_forward(space, _p(rgb), len(rgb), _p(out)):
        raise ValueError("orc_colorspace_forward")
    return out


def _colorspace_inverse(self, space, v):
    v = np.ascontiguousarray(v, np.float64).reshape(-1, 3)
    out = np.zeros_like(v)
    if self.lib.orc_colorspace_inverse(space, _p(v), len(v), _p(out)):
        raise ValueError("orc_colorspace_inverse")
    return out


def _ciecam_forward(self, xyz):
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    out = np.zeros_like(xyz)
    self.lib.orc_ciecam_forward(_p(xyz), len(xyz), _p(out))
    return out


def _ciecam_reverse(self, jch):
    jch = np.ascontiguousarray(jch, np.float64).reshape(-1, 3)
    out = np.zeros_like(jch)
    self.lib.orc_ciecam_reverse(_p(jch), len(jch), _p(out))
    return out


def _rgb_to_jch_level0(self, img, itype):
    """JCh pyramid level 0 of an RGB image: planar (3, h, w), int16 / int32."""
    img = np.ascontiguousarray(img, NP_I[itype])
    h, w, _ = img.shape
    out = np.zeros((3, h, w), np.int16 if itype == U8 else np.int32)
    if self.lib.orc_rgb_to_jch_level0(itype, w, h, _p(img), _p(out), w * h):
        raise ValueError("orc_rgb_to_jch_level0")
    return out


def _jch_level0_to_rgb_if(self, lv, mask, img, itype):
    lv = np.ascontiguousarray(lv, np.int16 if itype == U8 else np.int32)
    mask = np.ascontiguousarray(mask, np.uint8)
    out = np.array(img, NP_I[itype], copy=True)
    _, h, w = lv.shape
    if self.lib.orc_jch_level0_to_rgb_if(itype, w, h, _p(lv), w * h, _p(mask), _p(out)):
        raise ValueError("orc_jch_level0_to_rgb_if")
    return out


def _rgb_to_xyz_matrix(self):
    m = np.zeros(9, np.float64)
    self.lib.orc_rgb_to_xyz_matrix(_p(m))
    return m.reshape(3, 3)


def _nft_mask(self, white, black, norm=2):
    """(mask, rc): rc = -2 on excessive overlap (mask written anyway)."""
    white = np.ascontiguousarray(white, np.uint8)
    black = np.ascontiguousarray(black, np.uint8)
    h, w = white.shape
    mask = np.zeros((h, w), np.uint8)
    rc = self.lib.orc_nft_mask(_p(white), w, _p(black), w, w, h, norm, _p(mask), w)
    if rc not in (0, -2):
        raise ValueError(f"orc_nft_mask rc={rc}")
    return mask, rc


def _assemble(self, layers, itype, C, uw, uh, taken=None, one_at_a_time=False):
    """(placed, image (uh, uw, C), alpha (uh, uw), bb, taken) of one
    assemble() call over (samples, x, y, ftype) layers."""
    layers = [(np.ascontiguousarray(a), x, y, ft) for a, x, y, ft in layers]
    tab = layer_table(layers)
    tk = np.zeros(len(layers), np.int32) if taken is None else np.array(taken, np.int32)
    img = np.empty((uh, uw, C), NP_I[itype])
    al = np.empty((uh, uw), np.uint8)
    bb = np.zeros(4, np.int32)
    n = self.lib.orc_assemble(itype, C, uw, uh, len(layers), tab, int(one_at_a_time), _p(tk), _p(img), _p(al),
                              _p(bb))
    if n < 0:
        raise ValueError(f"orc_assemble rc={n}")
    return n, img, al, tuple(int(v) for v in bb), tk


def _inspect_overlap(self, a1, a2):
    a1 = np.ascontiguousarray(a1, np.uint8)
    a2 = np.ascontiguousarray(a2, np.uint8)
    h, w = a1.shape
    return self.lib.orc_inspect_overlap(_p(a1), w, _p(a2), w, w, h)


def _mask_bounds(self, mask):
    mask = np.ascontiguousarray(mask, np.uint8)
    h, w = mask.shape
    bb = np.zeros(4, np.int32)
    ok = self.lib.orc_mask_bounds(_p(mask), w, h, _p(bb))
    return bool(ok), tuple(int(v) for v in bb)


def _enblend_step(self, black, balpha, black_bb, white, walpha, white_bb, itype, wrap=False, exact_levels=0,
                  min_levels=1, norm=2, dither=1, seed=0):
    """One blending-loop iteration on union canvases: returns (rc, outcome,
    black, black_alpha, black_bb, info)."""
    uh, uw, C = black.shape
    b = np.array(black, NP_I[itype], copy=True)
    ba = np.array(balpha, np.uint8, copy=True)
    white = np.ascontiguousarray(white, NP_I[itype])
    walpha = np.ascontiguousarray(walpha, np.uint8)
    bbb = np.array(black_bb, np.int32)
    wbb = np.array(white_bb, np.int32)
    outcome = np.zeros(1, np.int32)
    info = np.zeros(9, np.int32)
    rc = self.lib.orc_enblend_step(itype, C, uw, uh, int(wrap), exact_levels, min_levels, norm, _p(b), _p(ba),
                                   _p(bbb), _p(white), _p(walpha), _p(wbb), dither, seed, _p(outcome), _p(info))
    return rc, int(outcome[0]), b, ba, tuple(int(v) for v in bbb), tuple(int(v) for v in info)


def _checkpoint(self, img, alpha, itype, otype):
    """(out (h, w, C + 1), component type) handed to the exporter."""
    h, w, C = img.shape
    img = np.ascontiguousarray(img, NP_I[itype])
    alpha = np.ascontiguousarray(alpha, np.uint8)
    out = np.zeros((h, w, C + 1), np.uint32 if itype == F32 and otype != F32 else NP_I[itype])
    ct = np.zeros(1, np.int32)
    rc = self.lib.orc_checkpoint(itype, C, w, h, _p(img), _p(alpha), otype, _p(out), _p(ct))
    if rc:
        raise ValueError(f"orc_checkpoint rc={rc}")
    return out, int(ct[0])


def _vigra_kernel(self, sigma, order):
    k = np.zeros(4096, np.float64)
    r = self.lib.orc_vigra_kernel(sigma, order, _p(k), len(k))
    if r < 0:
        raise ValueError("kernel too long")
    return k[:2 * r + 1].copy()


Oracle.vigra_kernel = _vigra_kernel
Oracle.assemble = _assemble
Oracle.inspect_overlap = _inspect_overlap
Oracle.mask_bounds = _mask_bounds
Oracle.enblend_step = _enblend_step
Oracle.euclid_1d = _euclid_1d
Oracle.set_blend_colorspace = _set_blend_colorspace
Oracle.rgb_to_blend_space = _rgb_to_blend_space
Oracle.blend_space_to_rgb_if = _blend_space_to_rgb_if
Oracle.colorspace_forward = _colorspace_forward
Oracle.colorspace_inverse = _colorspace_inverse
Oracle.rgb_to_xyz_matrix = _rgb_to_xyz_matrix
Oracle.ciecam_forward = _ciecam_forward
Oracle.ciecam_reverse = _ciecam_reverse
Oracle.rgb_to_jch_level0 = _rgb_to_jch_level0
Oracle.jch_level0_to_rgb_if = _jch_level0_to_rgb_if
Oracle.local_stddev_f32 = _local_stddev_f32
Oracle.checkpoint = _checkpoint
def _nft_mask_periodic(self, white, black, boundary, norm=2):
    """(mask, rc) with strip boundaries: 1 horizontal, 2 vertical, 3 double."""
    white = np.ascontiguousarray(white, np.uint8)
    black = np.ascontiguousarray(black, np.uint8)
    h, w = white.shape
    mask = np.zeros((h, w), np.uint8)
    rc = self.lib.orc_nft_mask_periodic(_p(white), w, _p(black), w, w, h, norm, boundary, _p(mask), w)
    if rc not in (0, -2):
        raise ValueError(f"orc_nft_mask_periodic rc={rc}")
    return mask, rc


Oracle.nft_mask_periodic = _nft_mask_periodic
Oracle.distance_transform = _distance_transform
Oracle.nft_mask = _nft_mask
Oracle.diff_forward = _diff_forward
Oracle.diff_backward = _diff_backward
Oracle.expand_T = _expand_T
Oracle.reduce_T = _reduce_T


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load_oracle():
    if not os.path.exists(ORACLE_LIB):
        build_oracle()
    lib = ctypes.CDLL(ORACLE_LIB)
    for name, args in _SIGS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    return Oracle(lib)
