// Pointwise stages of the pyramid path for gfx950: image <-> pyramid
// converters, the enblend level blend and alpha union.  One thread per pixel,
// rows mapped to blockIdx.y so that a wave touches one contiguous row segment
// of every plane.
#include <math.h>

#include "enbl_device.h"
#include "enbl_enfuse.h"
#include "enbl_internal.h"

#include <algorithm>

namespace enbl {
namespace {

constexpr int kTile = 256;

inline dim3 grid_for(int w, int h) { return dim3((w + kTile - 1) / kTile, h); }

// copyToPyramidImage with ConvertScalarToPyramidFunctor (src/fixmath.h:116-206):
// integer image -> (P)v << frac.
template <class I, class P, int C>
__global__ __launch_bounds__(kTile) void to_pyramid_int(const I* __restrict__ img, long ip, int w,
                                                        int frac, P* __restrict__ out, long pp, long pl)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const I* s = img + (long)y * ip + (long)x * C;
#pragma unroll
    for (int c = 0; c < C; ++c) out[c * pl + (long)y * pp + x] = (P)((P)s[c] << frac);
}

// float image -> double pyramid: v >= 0 ? 1 + log1p(v) : 1 / (1 - v)
// (src/fixmath.h:165-186).
template <int C>
__global__ __launch_bounds__(kTile) void to_pyramid_log(const float* __restrict__ img, long ip, int w,
                                                        double* __restrict__ out, long pp, long pl)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const float* s = img + (long)y * ip + (long)x * C;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const double v = (double)s[c];
        out[c * pl + (long)y * pp + x] = float_to_log_pyramid(v);
    }
}

// float -> fixed point: fromRealPromote(v * 2^frac) (src/fixmath.h:188-194).
template <class P, int C>
__global__ __launch_bounds__(kTile) void to_pyramid_fixed(const float* __restrict__ img, long ip, int w,
                                                          double scale, P* __restrict__ out, long pp, long pl)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const float* s = img + (long)y * ip + (long)x * C;
#pragma unroll
    for (int c = 0; c < C; ++c) out[c * pl + (long)y * pp + x] = from_real<P>((double)s[c] * scale);
}

// SpecializedConvertPyramidToScalarFunctor, integral pyramid -> integral
// image (src/fixmath.h:236-290), applied where mask != 0.
template <class P, class I, int C>
__global__ __launch_bounds__(kTile) void from_pyramid_int(const P* __restrict__ pyr, long pp, long pl, int w,
                                                          int frac, const uint8_t* __restrict__ mask, long mp,
                                                          I* __restrict__ img, long ip, int dither, uint64_t seed,
                                                          int imax)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    if (!mask[(long)y * mp + x]) return;
    const int32_t quarter = 1 << (frac - 2);
    const int32_t three_quarter = 3 << (frac - 2);
    const int32_t fmask = (1 << frac) - 1;
    I* d = img + (long)y * ip + (long)x * C;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int32_t v = pyr[c * pl + (long)y * pp + x];
        const int32_t q = v >> frac;
        const int32_t f = v & fmask;
        int32_t r;
        if (f >= quarter && f < three_quarter) {
            if (dither == ENBL_DITHER_DET) r = dither_rand(seed, x, y, c, frac) <= f ? q + 1 : q;
            else r = f >= (1 << (frac - 1)) ? q + 1 : q;
        } else if (f >= quarter) {
            r = q + 1;
        } else {
            r = q;
        }
        d[c] = (I)min(imax, max(0, r));
    }
}

// double pyramid -> float image: v >= 1 ? expm1(v - 1) : 1 - 1/v
// (src/fixmath.h:320-329), where mask != 0.
template <int C>
__global__ __launch_bounds__(kTile) void from_pyramid_log(const double* __restrict__ pyr, long pp, long pl, int w,
                                                          const uint8_t* __restrict__ mask, long mp,
                                                          float* __restrict__ img, long ip)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    if (!mask[(long)y * mp + x]) return;
    float* d = img + (long)y * ip + (long)x * C;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const double v = pyr[c * pl + (long)y * pp + x];
        d[c] = (float)log_pyramid_to_float(v);
    }
}

// CartesianBlendFunctor (src/blend.h:41-76) over one level.
template <class M, class P, int C>
__global__ __launch_bounds__(kTile) void blend_kernel(const M* __restrict__ mask, long mp,
                                                      const P* __restrict__ white, long wp, long wl,
                                                      P* __restrict__ black, long bp, long bl,
                                                      int w, double white_value)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const double wc = (double)mask[(long)y * mp + x] / white_value;
    if (wc >= 1.0) {
#pragma unroll
        for (int c = 0; c < C; ++c) black[c * bl + (long)y * bp + x] = white[c * wl + (long)y * wp + x];
        return;
    }
    if (wc <= 0.0) return;
    const double bc = 1.0 - wc;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        P* b = black + c * bl + (long)y * bp + x;
        const double v = (wc * (double)white[c * wl + (long)y * wp + x]) + (bc * (double)*b);
        *b = from_real<P>(v);
    }
}

// Zero fill (used instead of hipMemsetAsync so that captured graphs replay
// correctly).
__global__ __launch_bounds__(kTile) void fill_kernel(uint32_t* __restrict__ p, size_t words, uint8_t* tail, int ntail)
{
    const size_t i = (size_t)blockIdx.x * kTile + threadIdx.x;
    if (i < words) p[i] = 0u;
    if (i < (size_t)ntail) tail[i] = 0;
}

__global__ __launch_bounds__(kTile) void alpha_union_kernel(const uint8_t* __restrict__ s, long sp,
                                                            uint8_t* __restrict__ d, long dp, int w)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    if (s[(long)y * sp + x]) d[(long)y * dp + x] = 255;
}

template <class I, class P>
void to_int(int C, const void* img, long ip, int w, int h, int frac, void* out, long pp, long pl, hipStream_t s)
{
    if (C == 1) hipLaunchKernelGGL((to_pyramid_int<I, P, 1>), grid_for(w, h), dim3(kTile), 0, s, (const I*)img, ip, w, frac, (P*)out, pp, pl);
    else hipLaunchKernelGGL((to_pyramid_int<I, P, 3>), grid_for(w, h), dim3(kTile), 0, s, (const I*)img, ip, w, frac, (P*)out, pp, pl);
}

template <class P>
void to_fixed(int C, const void* img, long ip, int w, int h, int frac, void* out, long pp, long pl, hipStream_t s)
{
    const double scale = (double)(1u << frac);
    if (C == 1) hipLaunchKernelGGL((to_pyramid_fixed<P, 1>), grid_for(w, h), dim3(kTile), 0, s, (const float*)img, ip, w, scale, (P*)out, pp, pl);
    else hipLaunchKernelGGL((to_pyramid_fixed<P, 3>), grid_for(w, h), dim3(kTile), 0, s, (const float*)img, ip, w, scale, (P*)out, pp, pl);
}

template <class P, class I>
void from_int(int C, const void* pyr, long pp, long pl, int w, int h, int frac, const uint8_t* m, long mp,
              void* img, long ip, int dither, uint64_t seed, int imax, hipStream_t s)
{
    if (C == 1) hipLaunchKernelGGL((from_pyramid_int<P, I, 1>), grid_for(w, h), dim3(kTile), 0, s, (const P*)pyr, pp, pl, w, frac, m, mp, (I*)img, ip, dither, seed, imax);
    else hipLaunchKernelGGL((from_pyramid_int<P, I, 3>), grid_for(w, h), dim3(kTile), 0, s, (const P*)pyr, pp, pl, w, frac, m, mp, (I*)img, ip, dither, seed, imax);
}

template <class M, class P>
void blend_t(int C, int w, int h, const void* m, long mp, const void* wh, long wp, long wl,
             void* bk, long bp, long bl, double wv, hipStream_t s)
{
    if (C == 1) hipLaunchKernelGGL((blend_kernel<M, P, 1>), grid_for(w, h), dim3(kTile), 0, s, (const M*)m, mp, (const P*)wh, wp, wl, (P*)bk, bp, bl, w, wv);
    else hipLaunchKernelGGL((blend_kernel<M, P, 3>), grid_for(w, h), dim3(kTile), 0, s, (const M*)m, mp, (const P*)wh, wp, wl, (P*)bk, bp, bl, w, wv);
}

}  // namespace

int launch_to_pyramid(int itype, int ptype, int frac, int C, const void* img, long ip, int w, int h,
                      void* out, long pp, long pl, hipStream_t s)
{
    if (C != 1 && C != 3) return ENBL_E_UNSUPPORTED;
    if (itype == ENBL_U8 && ptype == ENBL_I16) { to_int<uint8_t, int16_t>(C, img, ip, w, h, frac, out, pp, pl, s); return 0; }
    if (itype == ENBL_U8 && ptype == ENBL_I32) { to_int<uint8_t, int32_t>(C, img, ip, w, h, frac, out, pp, pl, s); return 0; }
    if (itype == ENBL_U16 && ptype == ENBL_I32) { to_int<uint16_t, int32_t>(C, img, ip, w, h, frac, out, pp, pl, s); return 0; }
    if (itype == ENBL_F32 && ptype == ENBL_F64) {
        if (C == 1) hipLaunchKernelGGL((to_pyramid_log<1>), grid_for(w, h), dim3(kTile), 0, s, (const float*)img, ip, w, (double*)out, pp, pl);
        else hipLaunchKernelGGL((to_pyramid_log<3>), grid_for(w, h), dim3(kTile), 0, s, (const float*)img, ip, w, (double*)out, pp, pl);
        return 0;
    }
    if (itype == ENBL_F32 && ptype == ENBL_I32) { to_fixed<int32_t>(C, img, ip, w, h, frac, out, pp, pl, s); return 0; }
    if (itype == ENBL_F32 && ptype == ENBL_I16) { to_fixed<int16_t>(C, img, ip, w, h, frac, out, pp, pl, s); return 0; }
    return ENBL_E_UNSUPPORTED;
}

int launch_from_pyramid_if(int ptype, int itype, int frac, int C, const void* pyr, long pp, long pl,
                           int w, int h, const uint8_t* m, long mp, void* img, long ip,
                           int dither, uint64_t seed, hipStream_t s)
{
    if (C != 1 && C != 3) return ENBL_E_UNSUPPORTED;
    if (ptype == ENBL_I16 && itype == ENBL_U8) { from_int<int16_t, uint8_t>(C, pyr, pp, pl, w, h, frac, m, mp, img, ip, dither, seed, 255, s); return 0; }
    if (ptype == ENBL_I32 && itype == ENBL_U8) { from_int<int32_t, uint8_t>(C, pyr, pp, pl, w, h, frac, m, mp, img, ip, dither, seed, 255, s); return 0; }
    if (ptype == ENBL_I32 && itype == ENBL_U16) { from_int<int32_t, uint16_t>(C, pyr, pp, pl, w, h, frac, m, mp, img, ip, dither, seed, 65535, s); return 0; }
    if (ptype == ENBL_F64 && itype == ENBL_F32) {
        if (C == 1) hipLaunchKernelGGL((from_pyramid_log<1>), grid_for(w, h), dim3(kTile), 0, s, (const double*)pyr, pp, pl, w, m, mp, (float*)img, ip);
        else hipLaunchKernelGGL((from_pyramid_log<3>), grid_for(w, h), dim3(kTile), 0, s, (const double*)pyr, pp, pl, w, m, mp, (float*)img, ip);
        return 0;
    }
    return ENBL_E_UNSUPPORTED;
}

int launch_blend(int mtype, int ptype, int C, int w, int h, const void* m, long mp,
                 const void* wh, long wp, long wl, void* bk, long bp, long bl,
                 double wv, hipStream_t s)
{
    if (C != 1 && C != 3) return ENBL_E_UNSUPPORTED;
    if (mtype == ENBL_I16 && ptype == ENBL_I16) { blend_t<int16_t, int16_t>(C, w, h, m, mp, wh, wp, wl, bk, bp, bl, wv, s); return 0; }
    if (mtype == ENBL_I32 && ptype == ENBL_I32) { blend_t<int32_t, int32_t>(C, w, h, m, mp, wh, wp, wl, bk, bp, bl, wv, s); return 0; }
    if (mtype == ENBL_I32 && ptype == ENBL_F64) { blend_t<int32_t, double>(C, w, h, m, mp, wh, wp, wl, bk, bp, bl, wv, s); return 0; }
    if (mtype == ENBL_I16 && ptype == ENBL_I32) { blend_t<int16_t, int32_t>(C, w, h, m, mp, wh, wp, wl, bk, bp, bl, wv, s); return 0; }
    return ENBL_E_UNSUPPORTED;
}

int launch_fill(void* p, size_t bytes, hipStream_t s)
{
    // p is at least 4-byte aligned (device allocations and plane offsets).
    const size_t words = bytes / 4;
    const int ntail = (int)(bytes % 4);
    const size_t n = std::max(words, (size_t)ntail);
    if (n == 0) return 0;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + kTile - 1) / kTile)), dim3(kTile), 0, s, (uint32_t*)p, words,
                       (uint8_t*)p + words * 4, ntail);
    return 0;
}

int launch_alpha_union(const uint8_t* src, long sp, uint8_t* dst, long dp, int w, int h, hipStream_t s)
{
    hipLaunchKernelGGL(alpha_union_kernel, grid_for(w, h), dim3(kTile), 0, s, src, sp, dst, dp, w);
    return 0;
}

}  // namespace enbl
