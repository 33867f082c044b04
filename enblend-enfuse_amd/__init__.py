"""Python harness over the C ABI of the enblend/enfuse pyramid path on MI355X.

The product is the HIP shared library libenblend_amd.so (C ABI declared in
include/enblend_amd.h).  This module binds it with ctypes for the test and
benchmark harness; device buffers come from torch tensors (plumbing only).
There is no CPU fallback: loading fails loudly when the library is missing.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libenblend_amd.so")
HEADER = os.path.join(HERE, "..", "include", "enblend_amd.h")

ENBL_OK = 0
ENBL_E_PRECONDITION = -1
ENBL_E_SIZE = -2
ENBL_E_ARG = -3
ENBL_E_UNSUPPORTED = -4
ENBL_E_OOM = -5
ENBL_E_HIP = -6
ENBL_E_IO = -7

I16, I32, F64 = 0, 1, 2
U8, U16, F32, U32 = 10, 11, 12, 13
DITHER_NONE, DITHER_DET = 0, 1
CS_IDENTITY, CS_CIELAB, CS_CIELUV, CS_CIECAM = 0, 1, 2, 3
KERNELS_AUTO, KERNELS_DIRECT, KERNELS_STREAMING = 0, 1, 2

_c_int, _c_long, _vp = ctypes.c_int, ctypes.c_long, ctypes.c_void_p

_SIGS = {
    "enbl_context_create": (_c_int, [_c_int, ctypes.POINTER(_vp)]),
    "enbl_context_destroy": (_c_int, [_vp]),
    "enbl_context_set_stream": (_c_int, [_vp, _vp]),
    "enbl_context_get_stream": (_c_int, [_vp, ctypes.POINTER(_vp)]),
    "enbl_synchronize": (_c_int, [_vp]),
    "enbl_last_error": (ctypes.c_char_p, []),
    "enbl_profile_enable": (_c_int, [_vp, _c_int]),
    "enbl_profile_report": (_c_int, [_vp, ctypes.c_char_p, ctypes.c_size_t]),
    "enbl_distance_transform": (_c_int, [_vp, _c_int, _c_int, _vp, _c_long, _c_int, _vp, _c_long]),
    "enbl_nft_quotients": (_c_int, [_vp, _vp, _vp, _vp, _c_long]),
    "enbl_nft_mask": (_c_int, [_vp, _c_int, _c_int, _vp, _c_long, _vp, _c_long, _c_int,
                               _vp, _c_long]),
    "enbl_nft_mask_periodic": (_c_int, [_vp, _c_int, _c_int, _vp, _c_long, _vp, _c_long, _c_int, _c_int,
                                        _vp, _c_long]),
    "enbl_set_kernel_policy": (_c_int, [_c_int]),
    "enbl_set_strip_shape": (_c_int, [_c_int, _c_int]),
    "enbl_reduce": (_c_int, [_vp, _c_int, _c_int, _c_int,
                             _vp, _c_int, _c_int, _c_long, _c_long, _vp, _c_long,
                             _vp, _c_int, _c_int, _c_long, _c_long, _vp, _c_long]),
    "enbl_expand": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int,
                             _vp, _c_int, _c_int, _c_long, _c_long,
                             _vp, _c_int, _c_int, _c_long, _c_long]),
    "enbl_pyramid_create": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, ctypes.POINTER(_vp)]),
    "enbl_pyramid_destroy": (_c_int, [_vp]),
    "enbl_pyramid_levels": (_c_int, [_vp]),
    "enbl_pyramid_level": (_c_int, [_vp, _c_int, ctypes.POINTER(_vp), ctypes.POINTER(_c_int),
                                    ctypes.POINTER(_c_int), ctypes.POINTER(_c_long), ctypes.POINTER(_c_long)]),
    "enbl_pyramid_copy_level": (_c_int, [_vp, _c_int, _vp, _c_int]),
    "enbl_auto_levels": (_c_int, [_c_int, _c_int, _c_int, _c_int]),
    "enbl_copy_to_pyramid": (_c_int, [_vp, _c_int, _c_int, _vp, _c_long, _vp]),
    "enbl_set_blend_colorspace": (_c_int, [_vp, _c_int]),
    "enbl_rgb_to_blend_space": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _c_long, _vp, _c_long]),
    "enbl_rgb_to_jch": (_c_int, [_vp, _c_int, _c_int, _c_int, _vp, _c_long, _vp, _c_long, _c_long]),
    "enbl_jch_to_rgb_if": (_c_int, [_vp, _c_int, _c_int, _c_int, _vp, _c_long, _c_long, _vp, _c_long, _vp, _c_long]),
    "enbl_blend_space_to_rgb_if": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _c_long, _vp, _c_long, _vp,
                                            _c_long]),
    "enbl_gaussian_pyramid": (_c_int, [_vp, _c_int, _c_int, _c_int, _vp, _c_long, _vp, _c_long, _vp]),
    "enbl_gaussian_pyramid_noalloc": (_c_int, [_vp, _c_int, _c_int, _c_int, _vp, _c_long, _vp, _c_long, _vp]),
    "enbl_laplacian_pyramid": (_c_int, [_vp, _c_int, _c_int, _c_int, _vp, _c_long, _vp, _c_long, _vp]),
    "enbl_collapse_pyramid": (_c_int, [_vp, _c_int, _vp]),
    "enbl_blend": (_c_int, [_vp, _vp, _vp, _vp, ctypes.c_double]),
    "enbl_copy_from_pyramid_if": (_c_int, [_vp, _vp, _c_int, _c_int, _vp, _c_long, _vp, _c_int, _c_long,
                                           _c_int, ctypes.c_uint64]),
    "enbl_enfuse_mask": (_c_int, [_vp, _c_int, _c_int, _c_int, _vp, _c_long, _vp, _c_long, _vp, _vp, _c_long]),
    "enbl_enfuse": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int,
                             _vp, _c_long, _vp, _c_long, _c_int, ctypes.c_uint64]),
    "enbl_enfuse_batch": (_c_int, [_vp, _c_int, _c_int, _vp, _vp, _c_int, _c_int]),
    "enbl_batch_get_tuning": (_c_int, [_vp, _vp]),
    "enbl_batch_set_tuning": (_c_int, [_vp, _vp]),
    "enbl_batch_autotune": (_c_int, [_vp, _c_int, _c_int, _vp, _vp, _c_int, _c_int, _vp]),
    "enbl_diff_enfuse_forward": (_c_int, [_vp, _c_int, _c_int, _c_int, _vp, _vp, _c_int, _c_int, _vp]),
    "enbl_diff_enfuse_backward": (_c_int, [_vp, _c_int, _c_int, _c_int, _vp, _vp, _c_int, _c_int, _vp, _vp, _vp]),
    "enbl_expand_adjoint": (_c_int, [_vp, _vp, _c_long, _c_int, _c_int, _vp, _c_long, _c_int, _c_int]),
    "enbl_reduce_adjoint": (_c_int, [_vp, _vp, _c_long, _c_int, _c_int, _vp, _c_long, _c_int, _c_int]),
    "enbl_enblend_pair": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                                   _vp, _c_long, _vp, _c_long, _vp, _c_long, _vp, _c_long,
                                   _vp, _c_long, _c_int, ctypes.c_uint64]),
    "enbl_assemble": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _vp,
                               _vp, _c_long, _vp, _c_long, _vp]),
    "enbl_inspect_overlap": (_c_int, [_vp, _c_int, _c_int, _vp, _c_long, _vp, _c_long, _vp]),
    "enbl_mask_bounds": (_c_int, [_vp, _c_int, _c_int, _vp, _c_long, _vp, _vp]),
    "enbl_enblend_step": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                                   _vp, _c_long, _vp, _c_long, _vp, _vp, _c_long, _vp, _c_long, _vp,
                                   _c_int, ctypes.c_uint64, _vp]),
    "enbl_enblend_layers": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _c_int, _c_int,
                                     _c_int, _c_int, _c_int, ctypes.c_uint64, _vp, _c_long, _vp, _c_long, _vp,
                                     _vp]),
    "enbl_checkpoint": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _c_long, _vp, _c_long, _c_int, _vp,
                                 _c_long, _vp]),
    "enbl_tiff_info": (_c_int, [ctypes.c_char_p, _c_int, _vp]),
    "enbl_tiff_read": (_c_int, [ctypes.c_char_p, _c_int, _vp, _c_long]),
    "enbl_tiff_write": (_c_int, [ctypes.c_char_p, _c_int, _vp, _c_int]),
}


class Rect(ctypes.Structure):
    """enbl_rect: vigra::Rect2D [x0, x1) x [y0, y1)."""
    _fields_ = [("x0", _c_int), ("y0", _c_int), ("x1", _c_int), ("y1", _c_int)]

    def tuple(self):
        return (self.x0, self.y0, self.x1, self.y1)


class Layer(ctypes.Structure):
    """enbl_layer: one decoded input image on the union canvas."""
    _fields_ = [("data", _vp), ("pitch", _c_long), ("w", _c_int), ("h", _c_int), ("x", _c_int), ("y", _c_int),
                ("ftype", _c_int), ("bands", _c_int)]


def layer_table(layers):
    """ctypes array of Layer for (device tensor (h, w, bands), x, y, ftype)."""
    tab = (Layer * max(1, len(layers)))()
    for i, (t, x, y, ft) in enumerate(layers):
        h, w, bands = t.shape
        tab[i] = Layer(t.data_ptr(), w * bands, w, h, x, y, ft, bands)
    return tab

class TiffLayerInfo(ctypes.Structure):
    """enbl_tiff_layer_info."""
    _fields_ = [("w", _c_int), ("h", _c_int), ("bands", _c_int), ("extra", _c_int), ("ftype", _c_int),
                ("x", _c_int), ("y", _c_int), ("canvas_w", _c_int), ("canvas_h", _c_int),
                ("compression", _c_int), ("pages", _c_int)]


class TiffPage(ctypes.Structure):
    """enbl_tiff_page."""
    _fields_ = [("w", _c_int), ("h", _c_int), ("ftype", _c_int), ("bands", _c_int), ("extra", _c_int),
                ("data", _vp), ("pitch", _c_long), ("x", _c_int), ("y", _c_int), ("canvas_w", _c_int),
                ("canvas_h", _c_int)]


TIFF_NONE, TIFF_LZW, TIFF_DEFLATE, TIFF_PACKBITS = 1, 5, 8, 32773
_NP_OF_FTYPE = {U8: "uint8", U16: "uint16", U32: "uint32", F32: "float32"}


def _ftype_of(dtype):
    import numpy as np
    for k, v in _NP_OF_FTYPE.items():
        if np.dtype(v) == dtype:
            return k
    raise ValueError(f"no TIFF component type for {dtype}")


def tiff_info(path, page=0):
    """Header of one TIFF directory as a dict (enbl_tiff_info)."""
    info = TiffLayerInfo()
    check(lib().enbl_tiff_info(os.fsencode(path), page, ctypes.byref(info)))
    return {k: getattr(info, k) for k, _ in TiffLayerInfo._fields_}


def tiff_read(path, page=0):
    """(array (h, w, bands) on the host, info dict) of one TIFF directory."""
    import numpy as np
    info = tiff_info(path, page)
    arr = np.zeros((info["h"], info["w"], info["bands"]), dtype=_NP_OF_FTYPE[info["ftype"]])
    check(lib().enbl_tiff_read(os.fsencode(path), page, arr.ctypes.data, info["w"] * info["bands"]))
    return arr, info


def tiff_write(path, pages, compression=TIFF_NONE):
    """Write layers (array (h, w, bands), extra, x, y[, canvas_w, canvas_h]) as one TIFF."""
    import numpy as np
    keep = [np.ascontiguousarray(p[0]) for p in pages]
    tab = (TiffPage * len(pages))()
    for i, (p, a) in enumerate(zip(pages, keep)):
        if a.ndim == 2:
            a = keep[i] = a[:, :, None]
        h, w, bands = a.shape
        cw, ch = (p[4], p[5]) if len(p) > 4 else (0, 0)
        tab[i] = TiffPage(w, h, _ftype_of(a.dtype), bands, p[1], a.ctypes.data, w * bands, p[2], p[3], cw, ch)
    check(lib().enbl_tiff_write(os.fsencode(path), len(pages), tab, compression))


_lib = None

EXPOSURE_GAUSSIAN, EXPOSURE_LORENTZIAN, EXPOSURE_HALFSINE, EXPOSURE_FULLSINE, EXPOSURE_BISQUARE = range(5)
(GRAY_AVERAGE, GRAY_LSTAR, GRAY_PLSTAR, GRAY_LIGHTNESS, GRAY_VALUE, GRAY_ANTI_VALUE, GRAY_LUMINANCE,
 GRAY_MIXER) = range(8)


class EnfuseParams(ctypes.Structure):
    """enbl_enfuse_params (include/enblend_amd.h)."""
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
                  exposure_optimum=0.5, exposure_width=0.2, exposure_function=EXPOSURE_GAUSSIAN,
                  contrast_window=5, entropy_window=3, exposure_table=None, hard_mask=False,
                  edge_scale=0.0, lce_scale=0.0, lce_factor=0.0, min_curvature=0.0,
                  exposure_cutoff=None, entropy_cutoff=None, gray_projector=0, mixer=(0.0, 0.0, 0.0)):
    """Parameters with the reference's defaults (src/enfuse.cc:118-146)."""
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
        import numpy as np
        tab = np.ascontiguousarray(exposure_table, np.float64)
        p.exposure_table = tab.ctypes.data
        p._keep = tab
    return p


def enfuse(ctx, itype, imgs, alphas, prm, levels, out, out_alpha, wrap=False, dither=1, seed=0):
    """Fuse K device images (torch tensors (h, w, 3)) into out/out_alpha."""
    K = len(imgs)
    h, w = imgs[0].shape[0], imgs[0].shape[1]
    ip = (_vp * K)(*[t.data_ptr() for t in imgs])
    ap = (_vp * K)(*[t.data_ptr() for t in alphas])
    ipp = (_c_long * K)(*[w * 3] * K)
    app = (_c_long * K)(*[w] * K)
    check(lib().enbl_enfuse(ctx.h, itype, w, h, K, ip, ipp, ap, app, ctypes.byref(prm), levels, int(wrap),
                            _vp(out.data_ptr()), w * 3, _vp(out_alpha.data_ptr()), w, dither, seed))


class Stack(ctypes.Structure):
    """enbl_stack (include/enblend_amd.h)."""
    _fields_ = [("w", _c_int), ("h", _c_int), ("K", _c_int), ("levels", _c_int),
                ("imgs", _vp), ("ipitches", _vp), ("alphas", _vp), ("apitches", _vp),
                ("out", _vp), ("opitch", _c_long), ("out_alpha", _vp), ("oapitch", _c_long),
                ("seed", ctypes.c_uint64)]


class BatchTuning(ctypes.Structure):
    """enbl_batch_tuning (include/enblend_amd.h)."""
    _fields_ = [("rows", _c_int * 3), ("blocks_per_cu", _c_int * 3)]

    def as_dict(self):
        return {"rows": list(self.rows), "blocks_per_cu": list(self.blocks_per_cu)}


def _stack_array(stacks):
    """enbl_stack array for [(imgs, alphas, out, out_alpha, levels, seed), ...]
    of contiguous (h, w, 3) / (h, w) device tensors.  Returns (array, keep)."""
    arr = (Stack * len(stacks))()
    keep = []
    for i, (imgs, alphas, out, oa, levels, seed) in enumerate(stacks):
        K = len(imgs)
        h, w = imgs[0].shape[0], imgs[0].shape[1]
        ip = (_vp * K)(*[t.data_ptr() for t in imgs])
        ap = (_vp * K)(*[t.data_ptr() for t in alphas])
        ipp = (_c_long * K)(*[w * 3] * K)
        app = (_c_long * K)(*[w] * K)
        keep += [ip, ap, ipp, app]
        arr[i] = Stack(w, h, K, levels, ctypes.cast(ip, _vp), ctypes.cast(ipp, _vp), ctypes.cast(ap, _vp),
                       ctypes.cast(app, _vp), out.data_ptr(), w * 3, oa.data_ptr(), w, seed)
    return arr, keep


def enfuse_batch(ctx, itype, stacks, prm, wrap=False, dither=1):
    """Batched enfuse over [(imgs, alphas, out, out_alpha, levels, seed), ...]."""
    arr, keep = _stack_array(stacks)
    check(lib().enbl_enfuse_batch(ctx.h, itype, len(stacks), arr, ctypes.byref(prm), int(wrap), dither))
    del keep


def batch_tuning(ctx, tuning=None):
    """Get (tuning None) or set the batched passes' launch shape; returns it as a dict."""
    t = BatchTuning()
    if tuning is not None:
        t.rows[:] = list(tuning["rows"])
        t.blocks_per_cu[:] = list(tuning["blocks_per_cu"])
        check(lib().enbl_batch_set_tuning(ctx.h, ctypes.byref(t)))
    check(lib().enbl_batch_get_tuning(ctx.h, ctypes.byref(t)))
    return t.as_dict()


def batch_autotune(ctx, itype, stacks, prm, dither=1, repeats=2):
    """Run the autotuner on this batch; installs and returns the best shape."""
    arr, keep = _stack_array(stacks)
    t = BatchTuning()
    check(lib().enbl_batch_autotune(ctx.h, itype, len(stacks), arr, ctypes.byref(prm), dither, repeats, ctypes.byref(t)))
    del keep
    return t.as_dict()


class EnblError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"enblend_amd error {code}: {msg}")
        self.code = code


def library_path():
    return LIB_PATH


def lib():
    """Load libenblend_amd.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with enblend-enfuse_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != ENBL_OK:
        raise EnblError(rc, lib().enbl_last_error().decode())
    return rc


def ptr(t):
    """Device (or host) address of a torch tensor or None."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


class Context:
    """Owns an enbl_context on one HIP device."""

    def __init__(self, device=0):
        h = _vp()
        check(lib().enbl_context_create(device, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().enbl_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().enbl_synchronize(self.h))

    def stream(self):
        s = _vp()
        check(lib().enbl_context_get_stream(self.h, ctypes.byref(s)))
        return s.value

    def set_stream(self, hip_stream):
        """Order the library's work on an external HIP stream (None: its own)."""
        check(lib().enbl_context_set_stream(self.h, _vp(hip_stream) if hip_stream else None))

    def profile(self, on=True):
        check(lib().enbl_profile_enable(self.h, 1 if on else 0))

    def profile_report(self):
        buf = ctypes.create_string_buffer(1 << 16)
        check(lib().enbl_profile_report(self.h, buf, len(buf)))
        out = {}
        for line in buf.value.decode().splitlines():
            name, ms, n, nbytes, own = line.split()
            out[name] = {"ms": float(ms), "launches": int(n), "bytes": float(nbytes), "own_bytes": float(own)}
        return out


class Pyramid:
    """Device pyramid (enbl_pyramid); levels are read and written as numpy arrays."""

    def __init__(self, ctx, ptype, channels, w, h, levels):
        hnd = _vp()
        check(lib().enbl_pyramid_create(ctx.h, ptype, channels, w, h, levels, ctypes.byref(hnd)))
        self.h = hnd
        self.ctx = ctx
        self.ptype, self.channels, self.levels = ptype, channels, levels

    def close(self):
        if self.h:
            lib().enbl_pyramid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def level_info(self, l):
        d, w, h, p, pl = _vp(), _c_int(), _c_int(), _c_long(), _c_long()
        check(lib().enbl_pyramid_level(self.h, l, ctypes.byref(d), ctypes.byref(w), ctypes.byref(h),
                                       ctypes.byref(p), ctypes.byref(pl)))
        return d.value, w.value, h.value, p.value, pl.value

    def shape(self, l):
        _, w, h, _, _ = self.level_info(l)
        return (self.channels, h, w)

    def read(self, l):
        """Level l as a host numpy array of shape (channels, h, w)."""
        import numpy as np
        out = np.empty(self.shape(l), dtype=_NP[self.ptype])
        check(lib().enbl_pyramid_copy_level(self.h, l, out.ctypes.data_as(_vp), 0))
        return out

    def write(self, l, arr):
        """Upload a host array of shape (channels, h, w) into level l."""
        import numpy as np
        a = np.ascontiguousarray(arr, dtype=_NP[self.ptype])
        assert a.shape == self.shape(l), (a.shape, self.shape(l))
        check(lib().enbl_pyramid_copy_level(self.h, l, a.ctypes.data_as(_vp), 1))


_NP = {I16: "int16", I32: "int32", F64: "float64"}


def auto_levels(w, h, minimum_levels=1, exact_levels=0):
    return lib().enbl_auto_levels(w, h, minimum_levels, exact_levels)
