// TIFF layer files: the decode that vigra::importImageAlpha performs for
// enblend's import() (src/assemble.h:231-289) and the encode behind
// checkpoint()'s exportImageAlpha (src/assemble.h:130-224), as host code
// around the device pipeline (enbl_assemble, enbl_enblend_layers,
// enbl_checkpoint).
//
// Baseline TIFF 6.0 plus the extensions layer files use: 8-, 16- and 32-bit
// unsigned and 32-bit IEEE float samples, chunky or planar storage, strips
// or tiles, no compression, PackBits, LZW (with the TIFF code-width "early
// change") and Deflate (zlib), horizontal differencing (predictor 2) and the
// floating-point predictor (3, read only), either byte order, several
// directories (one layer each), ExtraSamples as alpha, and the layer's
// canvas position from XPosition/YPosition times the resolution, with the
// canvas size from the PIXAR_IMAGEFULLWIDTH/LENGTH tags that panorama
// stitchers write.  Files are written little-endian in strips.
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <zlib.h>

#include <algorithm>
#include <string>
#include <vector>

#include "enblend_amd.h"

namespace enbl {
namespace host {
int fail(int code, const char* fmt, ...);
}
}  // namespace enbl

using enbl::host::fail;

namespace {

enum : uint16_t {
    kWidth = 256, kLength = 257, kBits = 258, kCompression = 259, kPhotometric = 262, kStripOffsets = 273,
    kSamples = 277, kRowsPerStrip = 278, kStripBytes = 279, kXRes = 282, kYRes = 283, kPlanar = 284,
    kXPos = 286, kYPos = 287, kResUnit = 296, kPredictor = 317, kTileWidth = 322, kTileLength = 323,
    kTileOffsets = 324, kTileBytes = 325, kExtraSamples = 338, kSampleFormat = 339,
    kFullWidth = 33300, kFullLength = 33301
};
enum { kNone = 1, kLzw = 5, kDeflate = 8, kDeflateOld = 32946, kPackBits = 32773 };

constexpr double kWriteResolution = 150.0;   // dots per inch written with every layer

size_t type_size(uint16_t t)
{
    switch (t) {
    case 1: case 2: case 6: case 7: return 1;   // BYTE ASCII SBYTE UNDEFINED
    case 3: case 8: return 2;                   // SHORT SSHORT
    case 4: case 9: case 11: return 4;          // LONG SLONG FLOAT
    case 5: case 10: case 12: return 8;         // RATIONAL SRATIONAL DOUBLE
    default: return 0;
    }
}

struct Reader {
    std::vector<uint8_t> f;
    bool big = false;
    bool ok(size_t off, size_t n) const { return off <= f.size() && n <= f.size() - off; }
    uint16_t u16(size_t o) const {
        return big ? (uint16_t)(f[o] << 8 | f[o + 1]) : (uint16_t)(f[o] | f[o + 1] << 8);
    }
    uint32_t u32(size_t o) const {
        return big ? ((uint32_t)f[o] << 24 | (uint32_t)f[o + 1] << 16 | (uint32_t)f[o + 2] << 8 | f[o + 3])
                   : ((uint32_t)f[o] | (uint32_t)f[o + 1] << 8 | (uint32_t)f[o + 2] << 16 | (uint32_t)f[o + 3] << 24);
    }
};

struct Entry {
    uint16_t tag = 0, type = 0;
    uint32_t count = 0;
    size_t at = 0;   // file offset of the values
};

// A tag's value number i as a double (integers, rationals, floats).
bool value(const Reader& r, const Entry& e, uint32_t i, double* v)
{
    const size_t sz = type_size(e.type);
    if (!sz || i >= e.count || !r.ok(e.at + i * sz, sz)) return false;
    const size_t o = e.at + i * sz;
    switch (e.type) {
    case 1: case 7: *v = r.f[o]; return true;
    case 6: *v = (int8_t)r.f[o]; return true;
    case 3: *v = r.u16(o); return true;
    case 8: *v = (int16_t)r.u16(o); return true;
    case 4: *v = r.u32(o); return true;
    case 9: *v = (int32_t)r.u32(o); return true;
    case 5: { const uint32_t d = r.u32(o + 4); *v = d ? (double)r.u32(o) / d : 0.0; return true; }
    case 10: { const int32_t d = (int32_t)r.u32(o + 4); *v = d ? (double)(int32_t)r.u32(o) / d : 0.0; return true; }
    case 11: { uint32_t b = r.u32(o); float x; memcpy(&x, &b, 4); *v = x; return true; }
    case 12: {
        uint64_t b = r.big ? ((uint64_t)r.u32(o) << 32 | r.u32(o + 4)) : ((uint64_t)r.u32(o + 4) << 32 | r.u32(o));
        double x; memcpy(&x, &b, 8); *v = x; return true;
    }
    default: return false;
    }
}

struct Ifd {
    std::vector<Entry> e;
    const Entry* find(uint16_t tag) const {
        for (const Entry& x : e) if (x.tag == tag) return &x;
        return nullptr;
    }
};

int read_file(const char* path, Reader& r)
{
    FILE* fp = fopen(path, "rb");
    if (!fp) return fail(ENBL_E_IO, "cannot open \"%s\"", path);
    fseek(fp, 0, SEEK_END);
    const long n = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    r.f.resize(n > 0 ? (size_t)n : 0);
    const size_t got = n > 0 ? fread(r.f.data(), 1, (size_t)n, fp) : 0;
    fclose(fp);
    if (n <= 0 || got != (size_t)n) return fail(ENBL_E_IO, "cannot read \"%s\"", path);
    if (r.f.size() < 8) return fail(ENBL_E_IO, "\"%s\": not a TIFF file", path);
    if (r.f[0] == 'I' && r.f[1] == 'I') r.big = false;
    else if (r.f[0] == 'M' && r.f[1] == 'M') r.big = true;
    else return fail(ENBL_E_IO, "\"%s\": not a TIFF file", path);
    if (r.u16(2) == 43) return fail(ENBL_E_UNSUPPORTED, "\"%s\": BigTIFF is not supported", path);
    if (r.u16(2) != 42) return fail(ENBL_E_IO, "\"%s\": not a TIFF file", path);
    return ENBL_OK;
}

// The directories of the file (offsets checked, loops refused).
int directories(const Reader& r, const char* path, std::vector<Ifd>& out)
{
    uint32_t off = r.u32(4);
    std::vector<uint32_t> seen;
    while (off) {
        if (std::find(seen.begin(), seen.end(), off) != seen.end() || !r.ok(off, 2))
            return fail(ENBL_E_IO, "\"%s\": bad directory offset", path);
        seen.push_back(off);
        const uint16_t n = r.u16(off);
        if (!r.ok(off + 2, (size_t)n * 12 + 4)) return fail(ENBL_E_IO, "\"%s\": truncated directory", path);
        Ifd d;
        for (uint16_t i = 0; i < n; ++i) {
            const size_t o = off + 2 + (size_t)i * 12;
            Entry e;
            e.tag = r.u16(o);
            e.type = r.u16(o + 2);
            e.count = r.u32(o + 4);
            const size_t bytes = type_size(e.type) * (size_t)e.count;
            e.at = bytes <= 4 ? o + 8 : r.u32(o + 8);
            if (type_size(e.type) && !r.ok(e.at, bytes)) return fail(ENBL_E_IO, "\"%s\": tag %u out of the file", path, e.tag);
            d.e.push_back(e);
        }
        out.push_back(d);
        off = r.u32(off + 2 + (size_t)n * 12);
    }
    if (out.empty()) return fail(ENBL_E_IO, "\"%s\": no image directory", path);
    return ENBL_OK;
}

double tag_or(const Reader& r, const Ifd& d, uint16_t tag, double dflt, uint32_t i = 0)
{
    const Entry* e = d.find(tag);
    double v;
    if (e && value(r, *e, std::min(i, e->count ? e->count - 1 : 0), &v)) return v;
    return dflt;
}

struct Layout {
    int w = 0, h = 0, spp = 1, extra = 0, bits = 8, fmt = 1, ftype = 0, comp = kNone, pred = 1, planar = 1;
    bool tiled = false;
    int tw = 0, th = 0, rps = 0;
    int x = 0, y = 0, cw = 0, ch = 0;
};

int layout(const Reader& r, const Ifd& d, const char* path, Layout& L)
{
    L.w = (int)tag_or(r, d, kWidth, 0);
    L.h = (int)tag_or(r, d, kLength, 0);
    L.spp = (int)tag_or(r, d, kSamples, 1);
    L.bits = (int)tag_or(r, d, kBits, 1);
    for (int s = 1; s < L.spp; ++s)
        if ((int)tag_or(r, d, kBits, 1, (uint32_t)s) != L.bits)
            return fail(ENBL_E_UNSUPPORTED, "\"%s\": samples of different widths", path);
    L.fmt = (int)tag_or(r, d, kSampleFormat, 1);
    const Entry* ex = d.find(kExtraSamples);
    L.extra = ex ? (int)ex->count : 0;
    L.comp = (int)tag_or(r, d, kCompression, kNone);
    L.pred = (int)tag_or(r, d, kPredictor, 1);
    L.planar = (int)tag_or(r, d, kPlanar, 1);
    if (L.w < 1 || L.h < 1 || L.spp < 1 || L.spp > 16 || L.extra > L.spp)
        return fail(ENBL_E_IO, "\"%s\": bad image geometry", path);
    if (L.fmt == 1 && L.bits == 8) L.ftype = ENBL_U8;
    else if (L.fmt == 1 && L.bits == 16) L.ftype = ENBL_U16;
    else if (L.fmt == 1 && L.bits == 32) L.ftype = ENBL_U32;
    else if (L.fmt == 3 && L.bits == 32) L.ftype = ENBL_F32;
    else return fail(ENBL_E_UNSUPPORTED, "\"%s\": %d-bit samples of format %d", path, L.bits, L.fmt);
    if (L.comp != kNone && L.comp != kLzw && L.comp != kDeflate && L.comp != kDeflateOld && L.comp != kPackBits)
        return fail(ENBL_E_UNSUPPORTED, "\"%s\": compression %d", path, L.comp);
    if (L.pred != 1 && L.pred != 2 && L.pred != 3) return fail(ENBL_E_UNSUPPORTED, "\"%s\": predictor %d", path, L.pred);
    if (L.pred == 3 && L.fmt != 3) return fail(ENBL_E_UNSUPPORTED, "\"%s\": floating-point predictor on integers", path);
    if (L.planar != 1 && L.planar != 2) return fail(ENBL_E_IO, "\"%s\": bad planar configuration", path);
    L.tiled = d.find(kTileOffsets) != nullptr;
    if (L.tiled) {
        L.tw = (int)tag_or(r, d, kTileWidth, 0);
        L.th = (int)tag_or(r, d, kTileLength, 0);
        if (L.tw < 1 || L.th < 1) return fail(ENBL_E_IO, "\"%s\": bad tile size", path);
    } else {
        if (!d.find(kStripOffsets)) return fail(ENBL_E_IO, "\"%s\": no strips", path);
        const double rps = tag_or(r, d, kRowsPerStrip, 4294967295.0);
        L.rps = rps >= L.h ? L.h : (int)rps;
        if (L.rps < 1) return fail(ENBL_E_IO, "\"%s\": bad rows per strip", path);
    }
    // vigra's position: the offsets in resolution units times the resolution
    const double xr = tag_or(r, d, kXRes, 0.0), yr = tag_or(r, d, kYRes, 0.0);
    L.x = (int)lround(tag_or(r, d, kXPos, 0.0) * xr);
    L.y = (int)lround(tag_or(r, d, kYPos, 0.0) * yr);
    L.cw = (int)tag_or(r, d, kFullWidth, 0);
    L.ch = (int)tag_or(r, d, kFullLength, 0);
    return ENBL_OK;
}

// --- codecs ------------------------------------------------------------------

bool packbits_decode(const uint8_t* s, size_t n, uint8_t* d, size_t want)
{
    size_t i = 0, o = 0;
    while (o < want && i < n) {
        const int c = (int8_t)s[i++];
        if (c >= 0) {
            const size_t k = (size_t)c + 1;
            if (i + k > n || o + k > want) return false;
            memcpy(d + o, s + i, k);
            i += k;
            o += k;
        } else if (c != -128) {
            const size_t k = (size_t)(1 - c);
            if (i >= n || o + k > want) return false;
            memset(d + o, s[i++], k);
            o += k;
        }
    }
    return o == want;
}

void packbits_encode(const uint8_t* s, size_t n, std::vector<uint8_t>& out)
{
    size_t i = 0;
    while (i < n) {
        size_t run = 1;
        while (i + run < n && run < 128 && s[i + run] == s[i]) ++run;
        if (run >= 2) {
            out.push_back((uint8_t)(int8_t)(1 - (int)run));
            out.push_back(s[i]);
            i += run;
            continue;
        }
        size_t lit = 1;
        while (i + lit < n && lit < 128 && !(i + lit + 1 < n && s[i + lit] == s[i + lit + 1])) ++lit;
        out.push_back((uint8_t)(lit - 1));
        out.insert(out.end(), s + i, s + i + lit);
        i += lit;
    }
}

// TIFF LZW: 9..12-bit codes written MSB first, 256 = clear, 257 = end of
// information, the code width grows one code early (the decoder's table lags
// the encoder's by one entry).
bool lzw_decode(const uint8_t* s, size_t n, uint8_t* d, size_t want)
{
    std::vector<uint16_t> prefix(4096);
    std::vector<uint8_t> suffix(4096), first(4096);
    std::vector<uint16_t> len(4096);
    for (int c = 0; c < 256; ++c) { prefix[c] = 0xffff; suffix[c] = (uint8_t)c; first[c] = (uint8_t)c; len[c] = 1; }
    size_t bitpos = 0, o = 0;
    int width = 9, next = 258, old = -1;
    std::vector<uint8_t> tmp(4096);
    auto emit = [&](int code) -> bool {
        const int L = len[code];
        int c = code;
        for (int k = L - 1; k >= 0; --k) { tmp[k] = suffix[c]; c = prefix[c]; }
        const size_t take = std::min((size_t)L, want - o);
        memcpy(d + o, tmp.data(), take);
        o += take;
        return true;
    };
    while (o < want) {
        if (bitpos + width > n * 8) break;
        int code = 0;
        for (int b = 0; b < width; ++b, ++bitpos) code = code << 1 | ((s[bitpos >> 3] >> (7 - (bitpos & 7))) & 1);
        if (code == 257) break;
        if (code == 256) {
            width = 9;
            next = 258;
            old = -1;
            continue;
        }
        if (old < 0) {
            if (code > 255) return false;
            emit(code);
            old = code;
            continue;
        }
        if (code < next) {
            emit(code);
            if (next < 4096) {
                prefix[next] = (uint16_t)old; suffix[next] = first[code]; first[next] = first[old];
                len[next] = (uint16_t)(len[old] + 1);
                ++next;
            }
        } else if (code == next && next < 4096) {
            prefix[next] = (uint16_t)old; suffix[next] = first[old]; first[next] = first[old];
            len[next] = (uint16_t)(len[old] + 1);
            ++next;
            emit(code);
        } else {
            return false;
        }
        old = code;
        if (next >= (1 << width) - 1 && width < 12) ++width;
    }
    return o == want;
}

void lzw_encode(const uint8_t* s, size_t n, std::vector<uint8_t>& out)
{
    uint32_t acc = 0;
    int nacc = 0, width = 9;
    auto put = [&](int code) {
        acc = acc << width | (uint32_t)code;
        nacc += width;
        while (nacc >= 8) { out.push_back((uint8_t)(acc >> (nacc - 8))); nacc -= 8; }
        acc &= (1u << nacc) - 1u;
    };
    // dictionary: (prefix code << 8 | byte) -> code, open addressing
    const int kSlots = 1 << 13;
    std::vector<int32_t> key(kSlots), val(kSlots);
    auto reset = [&]() { std::fill(key.begin(), key.end(), -1); };
    reset();
    int next = 258;
    put(256);
    if (n == 0) { put(257); if (nacc) out.push_back((uint8_t)(acc << (8 - nacc))); return; }
    int ent = s[0];
    for (size_t i = 1; i < n; ++i) {
        const int c = s[i];
        const int32_t k = ent << 8 | c;
        uint32_t hh = ((uint32_t)k * 2654435761u) >> 19;
        int found = -1;
        while (key[hh] != -1) {
            if (key[hh] == k) { found = val[hh]; break; }
            hh = (hh + 1) & (kSlots - 1);
        }
        if (found >= 0) { ent = found; continue; }
        put(ent);
        key[hh] = k;
        val[hh] = next++;
        ent = c;
        if (next == 4094) {
            put(256);
            reset();
            next = 258;
            width = 9;
        } else if (next > (1 << width) - 1 && width < 12) {
            ++width;
        }
    }
    put(ent);
    ++next;   // the decoder adds an entry for the last code as well
    if (next > (1 << width) - 1 && width < 12) ++width;
    put(257);
    if (nacc) out.push_back((uint8_t)(acc << (8 - nacc)));
}

bool inflate_chunk(const uint8_t* s, size_t n, uint8_t* d, size_t want)
{
    uLongf got = (uLongf)want;
    const int rc = uncompress(d, &got, s, (uLong)n);
    return (rc == Z_OK || rc == Z_BUF_ERROR) && got == want;
}

// --- predictors ----------------------------------------------------------------

template <class T> void hor_acc(uint8_t* row, int w, int stride)
{
    T* p = reinterpret_cast<T*>(row);
    for (int i = stride; i < w * stride; ++i) p[i] = (T)(p[i] + p[i - stride]);
}
template <class T> void hor_diff(uint8_t* row, int w, int stride)
{
    T* p = reinterpret_cast<T*>(row);
    for (int i = w * stride - 1; i >= stride; --i) p[i] = (T)(p[i] - p[i - stride]);
}

void swab(uint8_t* p, size_t n, int size)
{
    if (size == 2) for (size_t i = 0; i + 1 < n; i += 2) std::swap(p[i], p[i + 1]);
    else if (size == 4) for (size_t i = 0; i + 3 < n; i += 4) { std::swap(p[i], p[i + 3]); std::swap(p[i + 1], p[i + 2]); }
}

// Undo the floating-point predictor of one row of w x stride 32-bit samples:
// byte differences, then the big-endian byte planes back into samples.
void fp_acc(uint8_t* row, int w, int stride)
{
    const size_t n = (size_t)w * stride * 4;
    for (size_t i = (size_t)stride; i < n; ++i) row[i] = (uint8_t)(row[i] + row[i - stride]);
    std::vector<uint8_t> t(row, row + n);
    const size_t cnt = (size_t)w * stride;
    for (size_t i = 0; i < cnt; ++i) {
        const uint32_t v = (uint32_t)t[i] << 24 | (uint32_t)t[cnt + i] << 16 | (uint32_t)t[2 * cnt + i] << 8 | t[3 * cnt + i];
        memcpy(row + 4 * i, &v, 4);
    }
}

// --- reading -------------------------------------------------------------------

int decode_chunk(const Reader& r, const Layout& L, size_t off, size_t nbytes, std::vector<uint8_t>& buf, size_t want,
                 const char* path)
{
    buf.assign(want, 0);
    if (!r.ok(off, nbytes)) return fail(ENBL_E_IO, "\"%s\": data out of the file", path);
    const uint8_t* s = r.f.data() + off;
    bool good;
    switch (L.comp) {
    case kNone: good = nbytes >= want; if (good) memcpy(buf.data(), s, want); break;
    case kPackBits: good = packbits_decode(s, nbytes, buf.data(), want); break;
    case kLzw: good = lzw_decode(s, nbytes, buf.data(), want); break;
    default: good = inflate_chunk(s, nbytes, buf.data(), want); break;
    }
    if (!good) return fail(ENBL_E_IO, "\"%s\": corrupt compressed data", path);
    return ENBL_OK;
}

// Samples of one decoded chunk row (cols x stride samples) into native order
// with the predictor undone.
void fix_row(const Reader& r, const Layout& L, uint8_t* row, int cols, int stride)
{
    const int sz = L.bits / 8;
    if (L.pred == 3) { fp_acc(row, cols, stride); return; }
    if (r.big && sz > 1) swab(row, (size_t)cols * stride * sz, sz);
    if (L.pred == 2) {
        if (sz == 1) hor_acc<uint8_t>(row, cols, stride);
        else if (sz == 2) hor_acc<uint16_t>(row, cols, stride);
        else hor_acc<uint32_t>(row, cols, stride);
    }
}

int read_image(const char* path, int page, Layout* out_layout, int* pages, void* buf, long pitch)
{
    Reader r;
    int rc = read_file(path, r);
    if (rc) return rc;
    std::vector<Ifd> dirs;
    rc = directories(r, path, dirs);
    if (rc) return rc;
    if (pages) *pages = (int)dirs.size();
    if (page < 0 || page >= (int)dirs.size()) return fail(ENBL_E_ARG, "\"%s\": no page %d", path, page);
    const Ifd& d = dirs[(size_t)page];
    Layout L;
    rc = layout(r, d, path, L);
    if (rc) return rc;
    if (out_layout) *out_layout = L;
    if (!buf) return ENBL_OK;
    if (pitch < (long)L.w * L.spp) return fail(ENBL_E_ARG, "enbl_tiff_read: pitch %ld below %d", pitch, L.w * L.spp);
    const int sz = L.bits / 8;
    uint8_t* dst = static_cast<uint8_t*>(buf);
    const size_t drow = (size_t)pitch * sz;
    const int planes = L.planar == 2 ? L.spp : 1;
    const int stride = L.planar == 2 ? 1 : L.spp;   // samples per pixel inside a chunk
    const Entry* offs = d.find(L.tiled ? kTileOffsets : kStripOffsets);
    const Entry* cnts = d.find(L.tiled ? kTileBytes : kStripBytes);
    const int cw = L.tiled ? L.tw : L.w, chh = L.tiled ? L.th : L.rps;
    const int across = (L.w + cw - 1) / cw, down = (L.h + chh - 1) / chh;
    const uint32_t nchunks = (uint32_t)across * down * planes;
    if (!offs || offs->count < nchunks) return fail(ENBL_E_IO, "\"%s\": missing data offsets", path);
    std::vector<uint8_t> chunk;
    for (int p = 0; p < planes; ++p) {
        for (int cy = 0; cy < down; ++cy) {
            for (int cx = 0; cx < across; ++cx) {
                const uint32_t idx = (uint32_t)(p * down + cy) * across + cx;
                double o = 0, nb = 0;
                value(r, *offs, idx, &o);
                if (cnts) value(r, *cnts, idx, &nb);
                const int rows = L.tiled ? L.th : std::min(chh, L.h - cy * chh);
                const size_t rowb = (size_t)cw * stride * sz;
                const size_t want = rowb * rows;
                if (!cnts) nb = (double)want;   // uncompressed files may omit the counts
                rc = decode_chunk(r, L, (size_t)o, (size_t)nb, chunk, want, path);
                if (rc) return rc;
                for (int yy = 0; yy < rows; ++yy) {
                    const int y = cy * chh + yy;
                    uint8_t* row = chunk.data() + rowb * yy;
                    fix_row(r, L, row, cw, stride);
                    if (y >= L.h) continue;
                    const int x0 = cx * cw, n = std::min(cw, L.w - x0);
                    uint8_t* dr = dst + drow * y;
                    if (L.planar == 1) {
                        memcpy(dr + (size_t)x0 * L.spp * sz, row, (size_t)n * L.spp * sz);
                    } else {
                        for (int x = 0; x < n; ++x)
                            memcpy(dr + ((size_t)(x0 + x) * L.spp + p) * sz, row + (size_t)x * sz, (size_t)sz);
                    }
                }
            }
        }
    }
    return ENBL_OK;
}

// --- writing ---------------------------------------------------------------------

struct Writer {
    std::vector<uint8_t> f;
    void u16(uint16_t v) { f.push_back((uint8_t)v); f.push_back((uint8_t)(v >> 8)); }
    void u32(uint32_t v) { for (int k = 0; k < 4; ++k) f.push_back((uint8_t)(v >> (8 * k))); }
    void put16(size_t at, uint16_t v) { f[at] = (uint8_t)v; f[at + 1] = (uint8_t)(v >> 8); }
    void put32(size_t at, uint32_t v) { for (int k = 0; k < 4; ++k) f[at + k] = (uint8_t)(v >> (8 * k)); }
    void align() { if (f.size() & 1) f.push_back(0); }
};

struct OutTag {
    uint16_t tag, type;
    std::vector<uint32_t> v;   // SHORT / LONG values, or (num, den) pairs of RATIONALs
};

int write_pages(const char* path, int npages, const enbl_tiff_page* pages, int compression)
{
    if (compression != kNone && compression != kLzw && compression != kDeflate && compression != kPackBits)
        return fail(ENBL_E_UNSUPPORTED, "enbl_tiff_write: compression %d", compression);
    Writer w;
    w.f.insert(w.f.end(), {'I', 'I', 42, 0});
    w.u32(0);
    size_t link = 4;   // where the next directory's offset goes
    for (int pg = 0; pg < npages; ++pg) {
        const enbl_tiff_page& P = pages[pg];
        int bits, fmt;
        switch (P.ftype) {
        case ENBL_U8: bits = 8; fmt = 1; break;
        case ENBL_U16: bits = 16; fmt = 1; break;
        case ENBL_U32: bits = 32; fmt = 1; break;
        case ENBL_F32: bits = 32; fmt = 3; break;
        default: return fail(ENBL_E_UNSUPPORTED, "enbl_tiff_write: pixel type %d", P.ftype);
        }
        if (P.w < 1 || P.h < 1 || P.bands < 1 || P.bands > 16 || P.extra < 0 || P.extra > P.bands - 1 || !P.data ||
            P.pitch < (long)P.w * P.bands || P.x < 0 || P.y < 0)
            return fail(ENBL_E_ARG, "enbl_tiff_write: bad page %d", pg);
        const int sz = bits / 8;
        const size_t rowb = (size_t)P.w * P.bands * sz;
        const int rps = (int)std::max<size_t>(1, std::min<size_t>((size_t)P.h, (256u << 10) / rowb));
        const int nstrips = (P.h + rps - 1) / rps;
        const bool pred = (compression == kLzw || compression == kDeflate) && fmt == 1;
        std::vector<uint32_t> soff, scnt;
        std::vector<uint8_t> raw, enc;
        const uint8_t* src = static_cast<const uint8_t*>(P.data);
        for (int s = 0; s < nstrips; ++s) {
            const int y0 = s * rps, rows = std::min(rps, P.h - y0);
            raw.resize(rowb * rows);
            for (int yy = 0; yy < rows; ++yy) {
                uint8_t* row = raw.data() + rowb * yy;
                memcpy(row, src + (size_t)P.pitch * sz * (y0 + yy), rowb);
                if (pred) {
                    if (sz == 1) hor_diff<uint8_t>(row, P.w, P.bands);
                    else if (sz == 2) hor_diff<uint16_t>(row, P.w, P.bands);
                    else hor_diff<uint32_t>(row, P.w, P.bands);
                }
            }
            enc.clear();
            if (compression == kNone) {
                enc = raw;
            } else if (compression == kPackBits) {
                for (int yy = 0; yy < rows; ++yy) packbits_encode(raw.data() + rowb * yy, rowb, enc);   // per row
            } else if (compression == kLzw) {
                lzw_encode(raw.data(), raw.size(), enc);
            } else {
                uLongf n = compressBound((uLong)raw.size());
                enc.resize(n);
                if (compress2(enc.data(), &n, raw.data(), (uLong)raw.size(), 6) != Z_OK)
                    return fail(ENBL_E_IO, "enbl_tiff_write: deflate failed");
                enc.resize(n);
            }
            w.align();
            if (w.f.size() + enc.size() > 0xffffffffu) return fail(ENBL_E_UNSUPPORTED, "enbl_tiff_write: file over 4 GiB");
            soff.push_back((uint32_t)w.f.size());
            scnt.push_back((uint32_t)enc.size());
            w.f.insert(w.f.end(), enc.begin(), enc.end());
        }
        const int colour = P.bands - P.extra;
        const uint32_t res = (uint32_t)kWriteResolution;
        std::vector<OutTag> tags;
        tags.push_back({kWidth, 4, {(uint32_t)P.w}});
        tags.push_back({kLength, 4, {(uint32_t)P.h}});
        tags.push_back({kBits, 3, std::vector<uint32_t>((size_t)P.bands, (uint32_t)bits)});
        tags.push_back({kCompression, 3, {(uint32_t)compression}});
        tags.push_back({kPhotometric, 3, {colour >= 3 ? 2u : 1u}});
        tags.push_back({kStripOffsets, 4, soff});
        tags.push_back({kSamples, 3, {(uint32_t)P.bands}});
        tags.push_back({kRowsPerStrip, 4, {(uint32_t)rps}});
        tags.push_back({kStripBytes, 4, scnt});
        tags.push_back({kXRes, 5, {res, 1}});
        tags.push_back({kYRes, 5, {res, 1}});
        tags.push_back({kPlanar, 3, {1}});
        tags.push_back({kXPos, 5, {(uint32_t)P.x, res}});
        tags.push_back({kYPos, 5, {(uint32_t)P.y, res}});
        tags.push_back({kResUnit, 3, {2}});
        if (pred) tags.push_back({kPredictor, 3, {2}});
        if (P.extra) tags.push_back({kExtraSamples, 3, std::vector<uint32_t>((size_t)P.extra, 2u)});   // unassociated alpha
        tags.push_back({kSampleFormat, 3, std::vector<uint32_t>((size_t)P.bands, (uint32_t)fmt)});
        if (P.canvas_w > 0 && P.canvas_h > 0) {
            tags.push_back({kFullWidth, 4, {(uint32_t)P.canvas_w}});
            tags.push_back({kFullLength, 4, {(uint32_t)P.canvas_h}});
        }
        // values longer than 4 bytes go before the directory
        std::vector<uint32_t> at(tags.size(), 0);
        for (size_t t = 0; t < tags.size(); ++t) {
            const OutTag& T = tags[t];
            const size_t cnt = T.type == 5 ? T.v.size() / 2 : T.v.size();
            const size_t bytes = cnt * type_size(T.type);
            if (bytes <= 4) continue;
            w.align();
            at[t] = (uint32_t)w.f.size();
            for (uint32_t v : T.v) {
                if (T.type == 3) w.u16((uint16_t)v);
                else w.u32(v);
            }
        }
        w.align();
        const uint32_t dir = (uint32_t)w.f.size();
        w.put32(link, dir);
        w.u16((uint16_t)tags.size());
        for (size_t t = 0; t < tags.size(); ++t) {
            const OutTag& T = tags[t];
            const size_t cnt = T.type == 5 ? T.v.size() / 2 : T.v.size();
            w.u16(T.tag);
            w.u16(T.type);
            w.u32((uint32_t)cnt);
            if (cnt * type_size(T.type) > 4) {
                w.u32(at[t]);
            } else if (T.type == 3) {
                w.u16(cnt > 0 ? (uint16_t)T.v[0] : 0);
                w.u16(cnt > 1 ? (uint16_t)T.v[1] : 0);
            } else {
                w.u32(T.v[0]);
            }
        }
        link = w.f.size();
        w.u32(0);
    }
    FILE* fp = fopen(path, "wb");
    if (!fp) return fail(ENBL_E_IO, "cannot create \"%s\"", path);
    const size_t put = fwrite(w.f.data(), 1, w.f.size(), fp);
    const int closed = fclose(fp);
    if (put != w.f.size() || closed != 0) return fail(ENBL_E_IO, "cannot write \"%s\"", path);
    return ENBL_OK;
}

}  // namespace

extern "C" {

int enbl_tiff_info(const char* path, int page, enbl_tiff_layer_info* info)
{
    if (!path || !info) return fail(ENBL_E_ARG, "enbl_tiff_info: null argument");
    Layout L;
    int pages = 0;
    const int rc = read_image(path, page, &L, &pages, nullptr, 0);
    if (rc) return rc;
    info->w = L.w;
    info->h = L.h;
    info->bands = L.spp;
    info->extra = L.extra;
    info->ftype = L.ftype;
    info->x = L.x;
    info->y = L.y;
    info->canvas_w = L.cw;
    info->canvas_h = L.ch;
    info->compression = L.comp;
    info->pages = pages;
    return ENBL_OK;
}

int enbl_tiff_read(const char* path, int page, void* buf, long pitch)
{
    if (!path || !buf) return fail(ENBL_E_ARG, "enbl_tiff_read: null argument");
    return read_image(path, page, nullptr, nullptr, buf, pitch);
}

int enbl_tiff_write(const char* path, int npages, const enbl_tiff_page* pages, int compression)
{
    if (!path || npages < 1 || !pages) return fail(ENBL_E_ARG, "enbl_tiff_write: bad argument");
    return write_pages(path, npages, pages, compression);
}

}  // extern "C"
