// CPU restatement of the enblend/enfuse pyramid path.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation and the CPU
// baseline of bench.py.  Nothing in the product links or loads this file.
//
// The reduce/expand functions restate the reference's single-pass SKIPSM
// state machines (Waltz & Miller, SPIE 1998) as a horizontal row filter
// followed by a per-column vertical recursion.  The two are equivalent to the
// reference's interleaved loops because each column's vertical state update
// only consumes that row's horizontal result.  Every arithmetic step keeps the
// reference's operation order so that the double-precision path is
// bit-identical (build with -ffp-contract=off).
//
// Integer accumulators use wrapping two's-complement arithmetic: the reference
// relies on Int32 SKIPSM sums (src/numerictraits.h:168-189) and x86-64 wraps on
// overflow; wrapping addition is order-independent, so any evaluation order of
// the same sum gives the same 32 bits.

#include "oracle_core.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif


using namespace orc;

extern "C" int orc_dither_rand(uint64_t seed, int x, int y, int c, int frac_bits)
{
    return orc::dither_draw(seed, x, y, c, frac_bits);
}



extern "C" {

int orc_reduce(int ptype, int wrap, const void* src, int sw, int sh, long spitch,
               const uint8_t* alpha, long apitch, void* dst, int dw, int dh, long dpitch,
               uint8_t* dalpha, long dapitch)
{
    if (alpha && !dalpha) return -3;
    switch (ptype) {
    case ORC_I16: return reduce_dispatch<int16_t>(wrap, (const int16_t*)src, sw, sh, spitch, alpha, apitch, (int16_t*)dst, dw, dh, dpitch, dalpha, dapitch);
    case ORC_I32: return reduce_dispatch<int32_t>(wrap, (const int32_t*)src, sw, sh, spitch, alpha, apitch, (int32_t*)dst, dw, dh, dpitch, dalpha, dapitch);
    case ORC_F64: return reduce_dispatch<double>(wrap, (const double*)src, sw, sh, spitch, alpha, apitch, (double*)dst, dw, dh, dpitch, dalpha, dapitch);
    }
    return -4;
}

int orc_expand(int ptype, int add, int wrap, const void* src, int sw, int sh, long spitch,
               void* dst, int dw, int dh, long dpitch)
{
    int rc = check_expand_sizes(sw, sh, dw, dh);
    if (rc) return rc;
    switch (ptype) {
    case ORC_I16: expand_t<int16_t>(add, wrap, (const int16_t*)src, sw, sh, spitch, (int16_t*)dst, dw, dh, dpitch); return 0;
    case ORC_I32: expand_t<int32_t>(add, wrap, (const int32_t*)src, sw, sh, spitch, (int32_t*)dst, dw, dh, dpitch); return 0;
    case ORC_F64: expand_t<double>(add, wrap, (const double*)src, sw, sh, spitch, (double*)dst, dw, dh, dpitch); return 0;
    }
    return -4;
}

int orc_auto_levels(int w, int h)
{
    unsigned s = (unsigned)std::min(w, h);
    unsigned levels = 1;
    while (levels <= 29) {
        if (s <= 8u) break;
        s = (s + 1u) >> 1;
        ++levels;
    }
    return (int)levels;
}

int orc_to_pyramid(int itype, int ptype, int frac, int C, const void* img, int w, int h, long ip,
                   void* planes, long pp, long pl)
{
    if (itype == ORC_U8 && ptype == ORC_I16) { to_pyramid_int<uint8_t, int16_t>(frac, C, (const uint8_t*)img, w, h, ip, (int16_t*)planes, pp, pl); return 0; }
    if (itype == ORC_U8 && ptype == ORC_I32) { to_pyramid_int<uint8_t, int32_t>(frac, C, (const uint8_t*)img, w, h, ip, (int32_t*)planes, pp, pl); return 0; }
    if (itype == ORC_U16 && ptype == ORC_I32) { to_pyramid_int<uint16_t, int32_t>(frac, C, (const uint16_t*)img, w, h, ip, (int32_t*)planes, pp, pl); return 0; }
    if (itype == ORC_F32 && ptype == ORC_F64) { to_pyramid_log(C, (const float*)img, w, h, ip, (double*)planes, pp, pl); return 0; }
    if (itype == ORC_F32 && ptype == ORC_I32) { to_pyramid_fixed<int32_t>(frac, C, (const float*)img, w, h, ip, (int32_t*)planes, pp, pl); return 0; }
    if (itype == ORC_F32 && ptype == ORC_I16) { to_pyramid_fixed<int16_t>(frac, C, (const float*)img, w, h, ip, (int16_t*)planes, pp, pl); return 0; }
    return -4;
}

int orc_from_pyramid_if(int ptype, int itype, int frac, int C, const void* planes, int w, int h,
                        long pp, long pl, const uint8_t* mask, long mp, void* img, long ip,
                        int dither, uint64_t seed)
{
    if (ptype == ORC_I16 && itype == ORC_U8) { from_pyramid_int<int16_t, uint8_t>(frac, C, (const int16_t*)planes, w, h, pp, pl, mask, mp, (uint8_t*)img, ip, dither, seed); return 0; }
    if (ptype == ORC_I32 && itype == ORC_U16) { from_pyramid_int<int32_t, uint16_t>(frac, C, (const int32_t*)planes, w, h, pp, pl, mask, mp, (uint16_t*)img, ip, dither, seed); return 0; }
    if (ptype == ORC_I32 && itype == ORC_U8) { from_pyramid_int<int32_t, uint8_t>(frac, C, (const int32_t*)planes, w, h, pp, pl, mask, mp, (uint8_t*)img, ip, dither, seed); return 0; }
    if (ptype == ORC_F64 && itype == ORC_F32) { from_pyramid_log(C, (const double*)planes, w, h, pp, pl, mask, mp, (float*)img, ip); return 0; }
    return -4;
}

int orc_blend_level(int mtype, int ptype, int C, int w, int h, const void* mask, long mp,
                    const void* white, long wp, long wl, void* black, long bp, long bl,
                    double white_value)
{
    if (mtype == ORC_I16 && ptype == ORC_I16) { blend_t<int16_t, int16_t>(C, w, h, (const int16_t*)mask, mp, (const int16_t*)white, wp, wl, (int16_t*)black, bp, bl, white_value); return 0; }
    if (mtype == ORC_I32 && ptype == ORC_I32) { blend_t<int32_t, int32_t>(C, w, h, (const int32_t*)mask, mp, (const int32_t*)white, wp, wl, (int32_t*)black, bp, bl, white_value); return 0; }
    if (mtype == ORC_I32 && ptype == ORC_F64) { blend_t<int32_t, double>(C, w, h, (const int32_t*)mask, mp, (const double*)white, wp, wl, (double*)black, bp, bl, white_value); return 0; }
    return -4;
}

int orc_enblend_pair(int itype, int C, int w, int h, int levels, int wrap,
                     const void* white, const uint8_t* walpha, void* black, uint8_t* balpha,
                     const uint8_t* mask, int dither, uint64_t seed)
{
    if (levels < 1 || levels > 29) return -1;
    const int cs = orc_blend_colorspace();
    if (C == 3 && cs == ORC_COLORSPACE_CIECAM && itype != ORC_U8 && itype != ORC_U16) return -4;
    if ((cs == ORC_COLORSPACE_CIELAB || cs == ORC_COLORSPACE_CIELUV) && C == 3) {
        // both images into the scaled blend space, the step there, the
        // result back to RGB under the union alpha
        if (itype != ORC_U8 && itype != ORC_U16) return -4;
        const size_t n = (size_t)w * h * 3 * (itype == ORC_U8 ? 1 : 2);
        std::vector<uint8_t> wl(n), bl(n);
        orc_rgb_to_blend_space(cs, itype, w, h, white, wl.data());
        orc_rgb_to_blend_space(cs, itype, w, h, black, bl.data());
        orc_set_blend_colorspace(ORC_COLORSPACE_IDENTITY);
        const int rc = orc_enblend_pair(itype, C, w, h, levels, wrap, wl.data(), walpha, bl.data(), balpha, mask,
                                        dither, seed);
        orc_set_blend_colorspace(cs);
        if (rc) return rc;
        return orc_blend_space_to_rgb_if(cs, itype, w, h, bl.data(), balpha, black);
    }
    int rc;
    if (itype == ORC_U8) {
        rc = enblend_pair_t<uint8_t, int16_t, int16_t>(C, w, h, levels, wrap, 7, 7, (const uint8_t*)white, walpha, (uint8_t*)black, balpha, mask, dither, seed);
    } else if (itype == ORC_U16) {
        rc = enblend_pair_t<uint16_t, int32_t, int32_t>(C, w, h, levels, wrap, 7, 15, (const uint16_t*)white, walpha, (uint16_t*)black, balpha, mask, dither, seed);
    } else if (itype == ORC_F32) {
        rc = enblend_pair_t<float, double, int32_t>(C, w, h, levels, wrap, 0, 15, (const float*)white, walpha, (float*)black, balpha, mask, dither, seed);
    } else {
        return -4;
    }
    return rc;
}

int orc_pyramid(int laplacian, int itype, int C, int w, int h, int levels, int wrap,
                const void* img, const uint8_t* alpha, void* out, long* n_out)
{
    if (levels < 1 || levels > 29) return -1;
    long total = 0;
    pyramid_dims(w, h, levels, C, &total);
    if (n_out) *n_out = total;
    if (itype == ORC_U8)
        return pyramid_t<uint8_t, int16_t>(laplacian, C, w, h, levels, wrap, 7, (const uint8_t*)img, alpha, (int16_t*)out);
    if (itype == ORC_U16)
        return pyramid_t<uint16_t, int32_t>(laplacian, C, w, h, levels, wrap, 7, (const uint16_t*)img, alpha, (int32_t*)out);
    if (itype == ORC_F32)
        return pyramid_t<float, double>(laplacian, C, w, h, levels, wrap, 0, (const float*)img, alpha, (double*)out);
    return -4;
}

int orc_collapse(int ptype, int C, int w, int h, int levels, int wrap, void* pyr)
{
    std::vector<Dims> d = pyramid_dims(w, h, levels, C, nullptr);
    switch (ptype) {
    case ORC_I16: collapse_levels<int16_t>((int16_t*)pyr, d, C, wrap); return 0;
    case ORC_I32: collapse_levels<int32_t>((int32_t*)pyr, d, C, wrap); return 0;
    case ORC_F64: collapse_levels<double>((double*)pyr, d, C, wrap); return 0;
    }
    return -4;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
