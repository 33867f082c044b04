"""TIFF layer files (enbl_tiff_info / enbl_tiff_read / enbl_tiff_write).

The reference reads and writes its layers through vigra's TIFF codec, that is
through libtiff (src/assemble.h:130-224, :231-289).  The codec here is pinned
three ways: round trips through itself over every sample type, band count,
compression, predictor and page layout; files written here read back by the
system's libtiff (loaded with ctypes, the library the reference links); and
files written by libtiff -- strips and tiles, chunky and planar, both byte
orders, predictors 1, 2 and 3 -- read here.  The libtiff legs are skipped
where that library is absent.
"""
import ctypes
import ctypes.util
import os

import numpy as np
import pytest

from conftest import load_package

enbl = load_package()

DTYPES = ["uint8", "uint16", "uint32", "float32"]
COMPRESSIONS = [enbl.TIFF_NONE, enbl.TIFF_LZW, enbl.TIFF_DEFLATE, enbl.TIFF_PACKBITS]


def _image(rng, shape, dtype):
    if dtype == "float32":
        a = (rng.standard_normal(shape) * 100).astype(np.float32)
    else:
        a = rng.integers(0, np.iinfo(dtype).max, shape, dtype=np.uint64, endpoint=True).astype(dtype)
    # long runs and smooth rows, so that every coder meets repeats and predictable data
    a[2:9, 1:-1] = a[2, 1]
    a[10:14] = np.cumsum(np.ones_like(a[10:14]), axis=1).astype(dtype)
    return a


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("compression", COMPRESSIONS)
@pytest.mark.parametrize("bands,extra", [(1, 0), (2, 1), (3, 0), (4, 1)])
def test_round_trip(tmp_path, dtype, compression, bands, extra):
    rng = np.random.default_rng(bands * 100 + compression % 97)
    a = _image(rng, (41, 67, bands), dtype)
    p = str(tmp_path / "l.tif")
    enbl.tiff_write(p, [(a, extra, 7, 13, 300, 200)], compression)
    b, info = enbl.tiff_read(p)
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    assert (info["w"], info["h"], info["bands"], info["extra"]) == (67, 41, bands, extra)
    assert (info["x"], info["y"], info["canvas_w"], info["canvas_h"]) == (7, 13, 300, 200)
    assert info["compression"] == compression and info["pages"] == 1


def test_lzw_long_input_resets_the_table(tmp_path):
    """Enough distinct strings to fill the 4094-entry table several times."""
    rng = np.random.default_rng(5)
    a = rng.integers(0, 256, (300, 700, 3), dtype=np.uint8)
    p = str(tmp_path / "big.tif")
    enbl.tiff_write(p, [(a, 0, 0, 0)], enbl.TIFF_LZW)
    b, _ = enbl.tiff_read(p)
    assert np.array_equal(a, b)


def test_pages_and_positions(tmp_path):
    rng = np.random.default_rng(1)
    layers = [(_image(rng, (30 + 5 * k, 50 - 3 * k, 4), "uint16"), 1, 10 * k, 3 * k, 400, 300) for k in range(3)]
    p = str(tmp_path / "multi.tif")
    enbl.tiff_write(p, layers, enbl.TIFF_DEFLATE)
    for k, (a, _, x, y, cw, ch) in enumerate(layers):
        b, info = enbl.tiff_read(p, k)
        assert np.array_equal(a, b)
        assert (info["x"], info["y"], info["pages"]) == (x, y, 3)


def test_errors(tmp_path):
    with pytest.raises(enbl.EnblError) as e:
        enbl.tiff_info(str(tmp_path / "missing.tif"))
    assert e.value.code == enbl.ENBL_E_IO
    bad = tmp_path / "bad.tif"
    bad.write_bytes(b"not a tiff at all")
    with pytest.raises(enbl.EnblError) as e:
        enbl.tiff_info(str(bad))
    assert e.value.code == enbl.ENBL_E_IO
    p = str(tmp_path / "one.tif")
    enbl.tiff_write(p, [(np.zeros((4, 4, 3), np.uint8), 0, 0, 0)])
    with pytest.raises(enbl.EnblError) as e:
        enbl.tiff_info(p, 1)
    assert e.value.code == enbl.ENBL_E_ARG
    # a truncated file is reported, not read past its end
    raw = open(p, "rb").read()
    (tmp_path / "cut.tif").write_bytes(raw[:40])
    with pytest.raises(enbl.EnblError):
        enbl.tiff_read(str(tmp_path / "cut.tif"))
    z = np.zeros((4, 4, 3), np.int16)
    page = enbl.TiffPage(4, 4, enbl.I16, 3, 0, z.ctypes.data, 12, 0, 0, 0, 0)
    assert enbl.lib().enbl_tiff_write(os.fsencode(str(tmp_path / "x.tif")), 1, ctypes.byref(page),
                                      enbl.TIFF_NONE) == enbl.ENBL_E_UNSUPPORTED
    with pytest.raises(enbl.EnblError) as e:
        enbl.tiff_write(str(tmp_path / "x.tif"), [(np.zeros((4, 4, 3), np.uint8), 0, -1, 0)])
    assert e.value.code == enbl.ENBL_E_ARG


# ---- libtiff, the library behind the reference's vigra codec -----------------

def _libtiff():
    for name in ("libtiff.so.5", "libtiff.so.6", ctypes.util.find_library("tiff")):
        if not name:
            continue
        try:
            return ctypes.CDLL(name)
        except OSError:
            continue
    return None


LT = _libtiff()
needs_libtiff = pytest.mark.skipif(LT is None, reason="the system libtiff is not loadable")

if LT is not None:
    LT.TIFFOpen.restype = ctypes.c_void_p
    LT.TIFFOpen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    LT.TIFFClose.argtypes = [ctypes.c_void_p]
    LT.TIFFReadScanline.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint16]
    LT.TIFFWriteScanline.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint16]
    LT.TIFFWriteTile.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_uint32, ctypes.c_uint16]
    LT.TIFFReadDirectory.argtypes = [ctypes.c_void_p]
    LT.TIFFWriteDirectory.argtypes = [ctypes.c_void_p]
    LT.TIFFSetErrorHandler.restype = ctypes.c_void_p
    LT.TIFFSetErrorHandler.argtypes = [ctypes.c_void_p]
    LT.TIFFSetWarningHandler.restype = ctypes.c_void_p
    LT.TIFFSetWarningHandler.argtypes = [ctypes.c_void_p]
    LT.TIFFSetWarningHandler(None)

U16, U32, DBL = ctypes.c_uint16, ctypes.c_uint32, ctypes.c_double
_BITS = {"uint8": (8, 1), "uint16": (16, 1), "uint32": (32, 1), "float32": (32, 3)}


def _set(tif, tag, *vals):
    assert LT.TIFFSetField(ctypes.c_void_p(tif), U32(tag), *vals) == 1, tag


def _get16(tif, tag):
    v = U16()
    assert LT.TIFFGetField(ctypes.c_void_p(tif), U32(tag), ctypes.byref(v)) == 1, tag
    return v.value


def _get32(tif, tag):
    v = U32()
    assert LT.TIFFGetField(ctypes.c_void_p(tif), U32(tag), ctypes.byref(v)) == 1, tag
    return v.value


def libtiff_read(path, page=0):
    tif = LT.TIFFOpen(os.fsencode(path), b"r")
    assert tif
    try:
        for _ in range(page):
            assert LT.TIFFReadDirectory(ctypes.c_void_p(tif)) == 1
        w, h = _get32(tif, 256), _get32(tif, 257)
        spp, bps = _get16(tif, 277), _get16(tif, 258)
        fmt = U16(1)
        LT.TIFFGetField(ctypes.c_void_p(tif), U32(339), ctypes.byref(fmt))
        dtype = {(8, 1): "uint8", (16, 1): "uint16", (32, 1): "uint32", (32, 3): "float32"}[(bps, fmt.value)]
        out = np.zeros((h, w, spp), dtype)
        for y in range(h):
            assert LT.TIFFReadScanline(ctypes.c_void_p(tif), out[y].ctypes.data, y, 0) == 1
        return out
    finally:
        LT.TIFFClose(ctypes.c_void_p(tif))


def libtiff_write(path, a, extra, compression, predictor=1, tiled=False, planar=1, big_endian=False):
    h, w, bands = a.shape
    bits, fmt = _BITS[str(a.dtype)]
    tif = LT.TIFFOpen(os.fsencode(path), b"wb" if big_endian else b"wl")
    assert tif
    try:
        _set(tif, 256, U32(w))
        _set(tif, 257, U32(h))
        _set(tif, 277, U16(bands))
        _set(tif, 258, U16(bits))
        _set(tif, 339, U16(fmt))
        _set(tif, 259, U16(compression))
        _set(tif, 262, U16(2 if bands - extra >= 3 else 1))
        _set(tif, 284, U16(planar))
        if predictor != 1:
            _set(tif, 317, U16(predictor))
        if extra:
            ex = (ctypes.c_uint16 * extra)(*([2] * extra))
            _set(tif, 338, U16(extra), ex)
        _set(tif, 282, DBL(72.0))
        _set(tif, 283, DBL(72.0))
        _set(tif, 296, U16(2))
        _set(tif, 286, DBL(5.0 / 72.0))
        _set(tif, 287, DBL(9.0 / 72.0))
        planes = [a] if planar == 1 else [a[:, :, s:s + 1] for s in range(bands)]
        if tiled:
            tw = th = 16
            _set(tif, 322, U32(tw))
            _set(tif, 323, U32(th))
            for s, pl in enumerate(planes):
                c = pl.shape[2]
                for ty in range(0, h, th):
                    for tx in range(0, w, tw):
                        tile = np.zeros((th, tw, c), a.dtype)
                        blk = pl[ty:ty + th, tx:tx + tw]
                        tile[:blk.shape[0], :blk.shape[1]] = blk
                        assert LT.TIFFWriteTile(ctypes.c_void_p(tif), tile.ctypes.data, tx, ty, 0, s) > 0
        else:
            _set(tif, 278, U32(5))
            for s, pl in enumerate(planes):
                for y in range(h):
                    row = pl[y].copy()   # libtiff's predictors work in place on the buffer
                    assert LT.TIFFWriteScanline(ctypes.c_void_p(tif), row.ctypes.data, y, s) == 1
    finally:
        LT.TIFFClose(ctypes.c_void_p(tif))


@needs_libtiff
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("compression", COMPRESSIONS)
def test_libtiff_reads_ours(tmp_path, dtype, compression):
    rng = np.random.default_rng(11)
    for bands, extra in ((3, 0), (4, 1), (2, 1)):
        a = _image(rng, (29, 45, bands), dtype)
        p = str(tmp_path / "ours.tif")
        enbl.tiff_write(p, [(a, extra, 0, 0), (a[::-1].copy(), extra, 0, 0)], compression)
        assert np.array_equal(libtiff_read(p).view(np.uint8), a.view(np.uint8))
        assert np.array_equal(libtiff_read(p, 1).view(np.uint8), a[::-1].view(np.uint8))


@needs_libtiff
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("compression", COMPRESSIONS)
@pytest.mark.parametrize("layout", ["strips", "tiles", "planar", "big-endian"])
def test_we_read_libtiff(tmp_path, dtype, compression, layout):
    rng = np.random.default_rng(12)
    predictors = [1]
    if compression in (enbl.TIFF_LZW, enbl.TIFF_DEFLATE) and not (dtype == "float32" and layout == "big-endian"):
        # (libtiff does not read back its own big-endian files with the
        # floating-point predictor, so that pair has no reference)
        predictors.append(3 if dtype == "float32" else 2)
    for pred in predictors:
        for bands, extra in ((3, 0), (4, 1)):
            a = _image(rng, (37, 50, bands), dtype)
            p = str(tmp_path / "lt.tif")
            libtiff_write(p, a, extra, compression, pred, tiled=layout == "tiles", planar=2 if layout == "planar" else 1,
                          big_endian=layout == "big-endian")
            b, info = enbl.tiff_read(p)
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (pred, bands)
            assert (info["extra"], info["x"], info["y"]) == (extra, 5, 9)
