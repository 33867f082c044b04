// CPU restatement of enblend's driver around the blend: image import with
// alpha thresholding, pre-assembly of non-overlapping images, the overlap
// inspection, the seam-mask bounding box, the ROI and level count, one
// iteration of the blending loop, and the checkpoint conversion for export.
// TEST INFRASTRUCTURE ONLY: checker of the HIP driver path.
//
//   import            src/assemble.h:231-289
//   assemble          src/assemble.h:300-444
//   checkpoint        src/assemble.h:130-224 (up to the vigra export call)
//   inspectOverlap    src/bounds.h:41-80
//   roiBounds         src/bounds.h:95-172
//   maskBounds        src/mask.h:520-590
//   enblendMain loop  src/enblend.h:127-545 (fine NFT mask, no optimiser)
//
// VIGRA semantics the path relies on (library not vendored; restated from its
// documented behaviour): Rect2D union/intersection/isEmpty/addBorder,
// Threshold(lower, upper, no, yes)(s) = (s < lower || upper < s) ? no : yes,
// linearRangeMapping(smin, smax, dmin, dmax): scale = (dmax - dmin) /
// (smax - smin) (1 when equal), offset = dmin / scale - smin, result =
// fromRealPromote(scale * (x + offset)) in the destination's RealPromote
// (double for integer types, float for float), FindBoundingRectangle.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <limits>
#include <type_traits>
#include <vector>

#include "oracle.h"
#include "oracle_core.h"

namespace orc {
// NumericTraits<UInt32>::fromRealPromote, with NaN as x86-64 converts it (0).
template <> inline uint32_t fromRealPromote<uint32_t>(double v)
{
    if (isnan(v)) return 0;
    if (v < 0.0) return 0;
    return v > 4294967295.0 ? 4294967295u : (uint32_t)(int64_t)(v + 0.5);
}
}  // namespace orc

namespace {

using orc::fromRealPromote;

struct Rect {   // vigra::Rect2D: [x0, x1) x [y0, y1)
    int x0, y0, x1, y1;
    bool empty() const { return x1 <= x0 || y1 <= y0; }
    int w() const { return x1 - x0; }
    int h() const { return y1 - y0; }
};

Rect rect_union(Rect a, const Rect& b) {
    if (b.empty()) return a;
    if (a.empty()) return b;
    a.x0 = std::min(a.x0, b.x0); a.y0 = std::min(a.y0, b.y0);
    a.x1 = std::max(a.x1, b.x1); a.y1 = std::max(a.y1, b.y1);
    return a;
}
Rect rect_intersect(Rect a, const Rect& b) {
    if (a.empty()) return a;
    if (b.empty()) return b;
    a.x0 = std::max(a.x0, b.x0); a.y0 = std::max(a.y0, b.y0);
    a.x1 = std::min(a.x1, b.x1); a.y1 = std::min(a.y1, b.y1);
    return a;
}

// rangeOfPixelType (src/common.h:578-608) for the file types handled here.
double range_max(int t) { return t == ORC_U8 ? 255.0 : t == ORC_U16 ? 65535.0 : 1.0; }

// A decoded sample of file type ftype as the image's component type T.
template <class T> T sample(const void* p, int ftype, long i) {
    if (ftype == ORC_U8) return (T)((const uint8_t*)p)[i];
    if (ftype == ORC_U16) return (T)((const uint16_t*)p)[i];
    return (T)((const float*)p)[i];
}

template <class T> struct Real { typedef double R; };
template <> struct Real<float> { typedef float R; };

template <class T> T from_real_t(double v) { return fromRealPromote<T>(v); }
template <> float from_real_t<float>(double v) { return (float)v; }

// import(): the image bands cast to T, the alpha band through
// Threshold<T, UInt8>(T(max/2), T(max), 0, 255) (or 255 everywhere without
// alpha), then linearRangeMapping from the file's range to T's range when
// they differ.
template <class T>
void import_t(int C, const orc_layer& L, T* img, uint8_t* alpha)
{
    const double fmax = range_max(L.ftype);
    const T lower = (T)(fmax / 2.0), upper = (T)fmax;
    const bool integral = !std::is_same<T, float>::value;
    const double tmin = 0.0;
    const double tmax = integral ? (double)std::numeric_limits<T>::max() : 1.0;
    const bool map = !(0.0 == tmin && fmax == tmax);
    typedef typename Real<T>::R M;
    // linearRangeMapping(T(0), T(fmax), T(tmin), T(tmax)), componentwise equal
    const T smin = (T)0.0, smax = (T)fmax, dmin = (T)tmin, dmax = (T)tmax;
    const M diff = (M)(smax - smin);
    const M scale = diff == (M)0 ? (M)1 : (M)(dmax - dmin) / diff;
    const M offset = (M)dmin / scale - (M)smin;
    for (int y = 0; y < L.h; ++y)
        for (int x = 0; x < L.w; ++x) {
            const long s = (long)y * L.pitch + (long)x * L.bands;
            for (int c = 0; c < C; ++c) {
                const T v = sample<T>(L.data, L.ftype, s + c);
                img[((long)y * L.w + x) * C + c] = map ? from_real_t<T>((double)(scale * ((M)v + offset))) : v;
            }
            uint8_t a = 255;
            if (L.bands > C) {
                const T v = sample<T>(L.data, L.ftype, s + C);
                a = (v < lower || upper < v) ? 0 : 255;
            }
            alpha[(long)y * L.w + x] = a;
        }
}

template <class T>
int assemble_t(int C, int uw, int uh, int n, const orc_layer* layers, int one_at_a_time, int* taken,
               T* image, uint8_t* alpha, int* bb)
{
    int first = -1;
    for (int i = 0; i < n; ++i)
        if (!taken[i]) { first = i; break; }
    if (first < 0) return 0;
    // new ImageType(inputUnion.size()): zero-initialised canvas
    std::fill(image, image + (size_t)uw * uh * C, (T)0);
    std::fill(alpha, alpha + (size_t)uw * uh, (uint8_t)0);
    int placed = 0;
    {
        const orc_layer& L = layers[first];
        std::vector<T> im((size_t)L.w * L.h * C);
        std::vector<uint8_t> a((size_t)L.w * L.h);
        import_t<T>(C, L, im.data(), a.data());
        for (int y = 0; y < L.h; ++y)
            for (int x = 0; x < L.w; ++x) {
                const long d = (long)(L.y + y) * uw + (L.x + x);
                for (int c = 0; c < C; ++c) image[d * C + c] = im[((long)y * L.w + x) * C + c];
                alpha[d] = a[(long)y * L.w + x];
            }
        taken[first] = 1;
        ++placed;
    }
    if (!one_at_a_time) {
        std::vector<int> accept;
        for (int i = first + 1; i < n; ++i) {
            if (taken[i]) continue;
            const orc_layer& L = layers[i];
            std::vector<T> im((size_t)L.w * L.h * C);
            std::vector<uint8_t> a((size_t)L.w * L.h);
            import_t<T>(C, L, im.data(), a.data());
            bool overlap = false;
            for (int y = 0; y < L.h && !overlap; ++y)
                for (int x = 0; x < L.w; ++x)
                    if (a[(long)y * L.w + x] && alpha[(long)(L.y + y) * uw + (L.x + x)]) { overlap = true; break; }
            if (overlap) continue;
            for (int y = 0; y < L.h; ++y)
                for (int x = 0; x < L.w; ++x) {
                    if (!a[(long)y * L.w + x]) continue;
                    const long d = (long)(L.y + y) * uw + (L.x + x);
                    for (int c = 0; c < C; ++c) image[d * C + c] = im[((long)y * L.w + x) * C + c];
                    alpha[d] = a[(long)y * L.w + x];
                }
            accept.push_back(i);
            ++placed;
        }
        // the used entries leave the list after the scan
        for (int i : accept) taken[i] = 1;
    }
    // FindBoundingRectangle over the canvas alpha
    bool valid = false;
    Rect r = {0, 0, 0, 0};
    for (int y = 0; y < uh; ++y)
        for (int x = 0; x < uw; ++x) {
            if (!alpha[(long)y * uw + x]) continue;
            if (!valid) { r = {x, y, x + 1, y + 1}; valid = true; continue; }
            r.x0 = std::min(r.x0, x); r.y0 = std::min(r.y0, y);
            r.x1 = std::max(r.x1, x + 1); r.y1 = std::max(r.y1, y + 1);
        }
    bb[0] = r.x0; bb[1] = r.y0; bb[2] = r.x1; bb[3] = r.y1;
    return placed;
}

bool layer_ok(int itype, int C, int uw, int uh, const orc_layer& L)
{
    if (L.bands != C && L.bands != C + 1) return false;
    if (L.x < 0 || L.y < 0 || L.w < 0 || L.h < 0 || L.x + L.w > uw || L.y + L.h > uh) return false;
    if (L.ftype == itype) return true;
    // narrower integer files into a wider image type
    return (L.ftype == ORC_U8 && (itype == ORC_U16 || itype == ORC_F32)) || (L.ftype == ORC_U16 && itype == ORC_F32);
}

// inspectOverlap over two alpha windows of one size: 0 none, 1 partial,
// 2 complete (enum Overlap, src/common.h:193).
int inspect_overlap(const uint8_t* a1, long p1, const uint8_t* a2, long p2, int w, int h)
{
    bool found_overlap = false, found_distinct2 = false;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const bool s1 = a1[(long)y * p1 + x] != 0, s2 = a2[(long)y * p2 + x] != 0;
            if (s1 && s2) found_overlap = true;
            else if (s2) found_distinct2 = true;
            if (found_overlap && found_distinct2) return 1;
        }
    return found_overlap ? 2 : 0;
}

// maskBounds: bounding box of the mask's transition pixels, relative to the
// mask; returns false when there is none.
bool mask_bounds(const uint8_t* m, int w, int h, Rect* out)
{
    Rect r = {w, h, 0, 0};
    auto add = [&](Rect q) { r = rect_union(r, q); };
    for (int x = 1; x < w; ++x)
        if (m[x - 1] != m[x]) add({x - 1, 0, x + 1, 1});
    for (int y = 1; y < h; ++y) {
        const uint8_t* row = m + (long)y * w;
        const uint8_t* up = row - w;
        if (up[0] != row[0]) add({0, y - 1, 1, y + 1});
        for (int x = 1; x < w; ++x) {
            if (row[x - 1] != row[x]) add({x - 1, y, x + 1, y + 1});
            if (up[x] != row[x]) add({x, y - 1, x + 1, y + 1});
        }
    }
    *out = r;
    return !r.empty();
}

// roiBounds: mBB grown by filterHalfWidth(MAX_PYRAMID_LEVELS = 29) =
// 2^30 - 4, made as wide as uBB's right edge when the seam mask wraps and
// the ROI leaves it, clipped to uBB, then the level count of the ROI's short
// side.
int roi_bounds(const Rect& mBB, const Rect& uBB, bool wrap_mask, int min_levels, int exact_levels, Rect* roi)
{
    const int border = (1 << 30) - 4;
    Rect r = {mBB.x0 - border, mBB.y0 - border, mBB.x1 + border, mBB.y1 + border};
    if (wrap_mask && (r.x0 < 0 || r.x1 > uBB.x1)) { r.x0 = 0; r.x1 = uBB.x1; }
    r = rect_intersect(r, uBB);
    *roi = r;
    unsigned s = (unsigned)std::min(r.w(), r.h());
    const unsigned lo = (unsigned)std::max(1, min_levels);
    unsigned levels = lo;
    while (levels <= 29) {
        if (s <= 8u) break;
        s = (s + 1u) >> 1;
        ++levels;
    }
    if (levels > lo) {
        if (exact_levels >= 1) levels = std::min(levels, (unsigned)exact_levels);
        else if (exact_levels < 0)
            levels = ((int)levels + exact_levels >= (int)lo) ? levels - (unsigned)(-exact_levels) : lo;
    }
    return (int)levels;
}

template <class T>
int step_t(int itype, int C, int uw, int uh, int wrap, int exact_levels, int min_levels, int norm,
           T* black, uint8_t* balpha, int* black_bb, const T* white, const uint8_t* walpha, const int* white_bb,
           int dither, uint64_t seed, int* outcome, int* info)
{
    const Rect bBB = {black_bb[0], black_bb[1], black_bb[2], black_bb[3]};
    const Rect wBB = {white_bb[0], white_bb[1], white_bb[2], white_bb[3]};
    const Rect uBB = rect_union(bBB, wBB);
    const long uoff = (long)uBB.y0 * uw + uBB.x0;
    const int ov = uBB.empty() ? 0 : inspect_overlap(balpha + uoff, uw, walpha + uoff, uw, uBB.w(), uBB.h());
    if (ov == 2) { *outcome = 1; return 0; }
    if (ov == 0 && exact_levels == 0) {
        for (long i = 0; i < (long)uw * uh; ++i) {
            if (!walpha[i]) continue;
            for (int c = 0; c < C; ++c) black[i * C + c] = white[i * C + c];
            balpha[i] = walpha[i];
        }
        black_bb[0] = uBB.x0; black_bb[1] = uBB.y0; black_bb[2] = uBB.x1; black_bb[3] = uBB.y1;
        *outcome = 2;
        return 0;
    }
    const bool wrap_mask = wrap && uBB.w() == uw;
    if (uBB.empty()) return -1;
    const int w = uBB.w(), h = uBB.h();
    std::vector<uint8_t> mask((size_t)w * h);
    int rc = wrap_mask ? orc_nft_mask_periodic(walpha + uoff, uw, balpha + uoff, uw, w, h, norm, 1, mask.data(), w)
                       : orc_nft_mask(walpha + uoff, uw, balpha + uoff, uw, w, h, norm, mask.data(), w);
    if (rc) return -1;
    Rect mBB;
    if (!mask_bounds(mask.data(), w, h, &mBB)) {
        if (mask[0] == 0) return -1;   // "mask is entirely black, but white image was not identified as redundant"
        mBB = uBB;
    } else {
        mBB = {mBB.x0 + uBB.x0, mBB.y0 + uBB.y0, mBB.x1 + uBB.x0, mBB.y1 + uBB.y0};
    }
    Rect roi;
    const int levels = roi_bounds(mBB, uBB, wrap_mask, min_levels, exact_levels, &roi);
    const int wrap_blend = wrap && roi.w() == uw;
    if (info) {
        info[0] = mBB.x0; info[1] = mBB.y0; info[2] = mBB.x1; info[3] = mBB.y1;
        info[4] = roi.x0; info[5] = roi.y0; info[6] = roi.x1; info[7] = roi.y1;
        info[8] = levels;
    }
    const int rw = roi.w(), rh = roi.h();
    // ROI windows of the canvases and of the mask (mask coordinates relative to uBB)
    std::vector<T> wr((size_t)rw * rh * C), br((size_t)rw * rh * C);
    std::vector<uint8_t> war((size_t)rw * rh), bar((size_t)rw * rh), mr((size_t)rw * rh);
    for (int y = 0; y < rh; ++y)
        for (int x = 0; x < rw; ++x) {
            const long s = (long)(roi.y0 + y) * uw + roi.x0 + x, d = (long)y * rw + x;
            for (int c = 0; c < C; ++c) { wr[d * C + c] = white[s * C + c]; br[d * C + c] = black[s * C + c]; }
            war[d] = walpha[s];
            bar[d] = balpha[s];
            mr[d] = mask[(long)(roi.y0 - uBB.y0 + y) * w + (roi.x0 - uBB.x0 + x)];
        }
    // the mask inside the ROI is cleared after its Gaussian pyramid is made;
    // white pixels under the remaining mask (outside the ROI) go to black
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const int X = uBB.x0 + x, Y = uBB.y0 + y;
            const bool in_roi = X >= roi.x0 && X < roi.x1 && Y >= roi.y0 && Y < roi.y1;
            if (in_roi || !mask[(long)y * w + x]) continue;
            for (int c = 0; c < C; ++c) black[((long)Y * uw + X) * C + c] = white[((long)Y * uw + X) * C + c];
        }
    // black alpha |= white alpha over whiteBB (outside the ROI here; the ROI
    // part is the union inside the pair step)
    for (int Y = std::max(wBB.y0, 0); Y < wBB.y1; ++Y)
        for (int X = std::max(wBB.x0, 0); X < wBB.x1; ++X) {
            const bool in_roi = X >= roi.x0 && X < roi.x1 && Y >= roi.y0 && Y < roi.y1;
            if (!in_roi && walpha[(long)Y * uw + X]) balpha[(long)Y * uw + X] = 255;
        }
    rc = orc_enblend_pair(itype, C, rw, rh, levels, wrap_blend, wr.data(), war.data(), br.data(), bar.data(),
                          mr.data(), dither, seed);
    if (rc) return rc;
    for (int y = 0; y < rh; ++y)
        for (int x = 0; x < rw; ++x) {
            const long d = (long)(roi.y0 + y) * uw + roi.x0 + x, s = (long)y * rw + x;
            for (int c = 0; c < C; ++c) black[d * C + c] = br[s * C + c];
            balpha[d] = bar[s];
        }
    black_bb[0] = uBB.x0; black_bb[1] = uBB.y0; black_bb[2] = uBB.x1; black_bb[3] = uBB.y1;
    *outcome = 0;
    return 0;
}

// checkpoint(): the image handed to the exporter (identity, narrowing in the
// image's own type, or float rescaled into UInt32) and the alpha through
// Threshold(0, 0, max, 0), i.e. alpha != 0 -> AlphaTraits<D>::max().
template <class T, class D>
void checkpoint_t(int C, int w, int h, const T* img, const uint8_t* alpha, bool map, double smin, double smax,
                  double dmin, double dmax, D amax, D* out)
{
    // linearRangeMapping(T(smin), T(smax), D(dmin), D(dmax)) with a double
    // multiplier (the destination types here are integral)
    const double diff = (double)((T)smax - (T)smin);
    const double scale = diff == 0.0 ? 1.0 : (double)((D)dmax - (D)dmin) / diff;
    const double offset = (double)(D)dmin / scale - (double)(T)smin;
    for (long i = 0; i < (long)w * h; ++i) {
        for (int c = 0; c < C; ++c) {
            const T v = img[i * C + c];
            out[i * (C + 1) + c] = map ? fromRealPromote<D>(scale * ((double)v + offset)) : (D)v;
        }
        out[i * (C + 1) + C] = alpha[i] ? amax : (D)0;
    }
}

}  // namespace

extern "C" int orc_assemble(int itype, int C, int uw, int uh, int n, const orc_layer* layers, int one_at_a_time,
                            int* taken, void* image, uint8_t* alpha, int* bb)
{
    if (C != 1 && C != 3) return -4;
    for (int i = 0; i < n; ++i)
        if (!layer_ok(itype, C, uw, uh, layers[i])) return -4;
    if (itype == ORC_U8) return assemble_t<uint8_t>(C, uw, uh, n, layers, one_at_a_time, taken, (uint8_t*)image, alpha, bb);
    if (itype == ORC_U16) return assemble_t<uint16_t>(C, uw, uh, n, layers, one_at_a_time, taken, (uint16_t*)image, alpha, bb);
    if (itype == ORC_F32) return assemble_t<float>(C, uw, uh, n, layers, one_at_a_time, taken, (float*)image, alpha, bb);
    return -4;
}

extern "C" int orc_inspect_overlap(const uint8_t* a1, long p1, const uint8_t* a2, long p2, int w, int h)
{
    return inspect_overlap(a1, p1, a2, p2, w, h);
}

extern "C" int orc_mask_bounds(const uint8_t* mask, int w, int h, int* bb)
{
    Rect r;
    const bool ok = mask_bounds(mask, w, h, &r);
    bb[0] = r.x0; bb[1] = r.y0; bb[2] = r.x1; bb[3] = r.y1;
    return ok ? 1 : 0;
}

extern "C" int orc_enblend_step(int itype, int C, int uw, int uh, int wrap, int exact_levels, int min_levels,
                                int norm, void* black, uint8_t* balpha, int* black_bb, const void* white,
                                const uint8_t* walpha, const int* white_bb, int dither, uint64_t seed,
                                int* outcome, int* info)
{
    if (C != 1 && C != 3) return -4;
    if (itype == ORC_U8)
        return step_t<uint8_t>(itype, C, uw, uh, wrap, exact_levels, min_levels, norm, (uint8_t*)black, balpha,
                               black_bb, (const uint8_t*)white, walpha, white_bb, dither, seed, outcome, info);
    if (itype == ORC_U16)
        return step_t<uint16_t>(itype, C, uw, uh, wrap, exact_levels, min_levels, norm, (uint16_t*)black, balpha,
                                black_bb, (const uint16_t*)white, walpha, white_bb, dither, seed, outcome, info);
    if (itype == ORC_F32)
        return step_t<float>(itype, C, uw, uh, wrap, exact_levels, min_levels, norm, (float*)black, balpha,
                             black_bb, (const float*)white, walpha, white_bb, dither, seed, outcome, info);
    return -4;
}

extern "C" int orc_checkpoint(int itype, int C, int w, int h, const void* img, const uint8_t* alpha, int otype,
                              void* out, int* out_ctype)
{
    if (C != 1 && C != 3) return -4;
    const double imin = 0.0, imax = itype == ORC_U8 ? 255.0 : itype == ORC_U16 ? 65535.0 : 1.0;
    const double omin = 0.0, omax = range_max(otype);
    if (otype != ORC_U8 && otype != ORC_U16 && otype != ORC_F32) return -4;
    if (imin <= omin && imax >= omax) {
        const bool map = !(imin == omin && imax == omax);
        *out_ctype = itype;
        if (itype == ORC_U8)
            checkpoint_t<uint8_t, uint8_t>(C, w, h, (const uint8_t*)img, alpha, map, imin, imax, omin, omax, 255, (uint8_t*)out);
        else if (itype == ORC_U16)
            checkpoint_t<uint16_t, uint16_t>(C, w, h, (const uint16_t*)img, alpha, map, imin, imax, omin, omax, 65535, (uint16_t*)out);
        else
            checkpoint_t<float, float>(C, w, h, (const float*)img, alpha, false, imin, imax, omin, omax, 1.0f, (float*)out);
        return 0;
    }
    // rescale into IntegralSelect<ImagePixelType>: UInt32 for float images;
    // an integer image keeps its own type, whose cast of the output maximum
    // overflows (8-bit images written as 16-bit): not provided
    if (itype != ORC_F32) return -4;
    *out_ctype = ORC_U32;
    checkpoint_t<float, uint32_t>(C, w, h, (const float*)img, alpha, true, imin, imax, omin, omax, 4294967295u,
                                  (uint32_t*)out);
    return 0;
}
