// Batched enfuse passes for gfx950: many exposure stacks of uneven size and
// depth, one persistent launch per dependency level.
//
// Reference: enfuseMain (src/enfuse.h:1217-1609) with its exposure and
// saturation functors (src/enfuse.h:622-780), the pyramid reduce/expand of
// src/pyramid.h and the collapse of src/pyramid.h:1837-1886.  Every output
// value is computed by the plain per-stack path's (and so the reference's)
// operations in the same order; what changes is the schedule:
//   * one weights pass computes all K weights of a pixel, their norm, the
//     union alpha and the K normalised fixed-point mask levels 0;
//   * per level, one reduce pass builds the next level of every image and
//     mask pyramid of every stack;
//   * per level, from the coarsest down, one pass evaluates each image's
//     Laplacian value on the fly (G_m[l] - expand(G_m[l+1])), multiplies it
//     by its mask, accumulates the K images in the reference's order and
//     adds expand(result[l+1]) - the merge and the collapse step in one.
//     No Laplacian level and no uncollapsed result level is ever stored;
//     level 0 goes straight into the dithered output image.
// A launch is a persistent grid whose blocks walk a packed list of 64-column
// tiles taken from all jobs of the pass (binary search on the tile prefix).
#include <math.h>
#include <type_traits>

#include "enbl_batch.h"
#include "enbl_device.h"
#include "enbl_internal.h"
#include "pyr_common.h"
#include "reduce_stream.h"

namespace enbl {
namespace {

__device__ __forceinline__ int find_job(const int* __restrict__ tile0, int njobs, int t)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile0[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <class T> struct ImgMax;
template <> struct ImgMax<uint8_t> { static constexpr int i = 255; static constexpr double d = 255.0; };
template <> struct ImgMax<uint16_t> { static constexpr int i = 65535; static constexpr double d = 65535.0; };

// Level samples: a pyramid plane, or the interleaved integer image converted
// by ConvertScalarToPyramidFunctor ((P)v << frac, src/fixmath.h:116-163).
template <class P> struct PlaneLoad {
    typedef typename Pix<P>::S S;
    __device__ static __forceinline__ S get(const BPlane& b, int y, int x, int ch, int) {
        return Pix<P>::load(*gmem(static_cast<const P*>(b.p) + ch * b.plane + (long)y * b.pitch + x));
    }
};
template <class I, class P> struct ImageLoad {
    typedef typename Pix<P>::S S;
    __device__ static __forceinline__ S get(const BPlane& b, int y, int x, int ch, int frac) {
        return (S)(P)((P)*gmem(static_cast<const I*>(b.p) + (long)y * b.pitch + 3L * x + ch) << frac);
    }
};
// Float images: the log transform of src/fixmath.h:165-186 in double.
template <> struct ImageLoad<float, double> {
    __device__ static __forceinline__ double get(const BPlane& b, int y, int x, int ch, int) {
        return float_to_log_pyramid((double)*gmem(static_cast<const float*>(b.p) + (long)y * b.pitch + 3L * x + ch));
    }
};

// "average" grayscale projector (src/mga.h:188-193) with the integer
// fromRealPromote of the image type: (int)((p0 + p1 + p2) / 3.0 + 0.5).  For
// an integer sum S the quotient S / 3 is never within 1/6 of a half, so the
// double form equals (S + 1) / 3 exactly (checked over every sum of 8- and
// 16-bit samples by the test suite's exhaustive comparison).
template <class T> __device__ __forceinline__ int gray_code(T p0, T p1, T p2)
{
    return ((int)p0 + (int)p1 + (int)p2 + 1) / 3;
}

// ---------------------------------------------------------------------------
// Weights: exposure (host table by gray code) then saturation, float sums in
// the reference's order (src/enfuse.h:1000-1008, :1131-1135); norm += w in
// image order (:1377-1385); normalisation and fixed-point mask level 0
// (:1529-1546, src/fixmath.h:188-194).
// ---------------------------------------------------------------------------
// The weights of a pixel pair (x, x+1) for every image of the stack, all
// loads issued before any arithmetic; NK bounds the stack depth this
// instance keeps in registers (deeper stacks take the looping instance).
// The raw samples of a pixel pair (x, x+1) of an interleaved RGB image,
// packed: two 16-bit RGB pixels are the pair's 12 bytes (one dwordx3 load
// when aligned), two 8-bit ones six bytes in d[0] and the low half of d[1].
// Keeping the words packed until the weights are formed holds a deep
// stack's loads in flight with few registers.
template <class T>
__device__ __forceinline__ void load_raw(const T* p, bool two, uint32_t (&d)[3])
{
    if constexpr (sizeof(T) == 2) {
        if (two && ((uintptr_t)p & 3) == 0) {
            const auto* g = gmem(reinterpret_cast<const uint32_t*>(p));
            d[0] = g[0]; d[1] = g[1]; d[2] = g[2];
            return;
        }
        const auto* g = gmem(p);
        uint32_t v[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] = (k < 3 || two) ? (uint32_t)g[k] : 0u;
        d[0] = v[0] | (v[1] << 16); d[1] = v[2] | (v[3] << 16); d[2] = v[4] | (v[5] << 16);
    } else {
        const auto* g = gmem(p);
        uint32_t v[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] = (k < 3 || two) ? (uint32_t)g[k] : 0u;
        d[0] = v[0] | (v[1] << 8) | (v[2] << 16) | (v[3] << 24);
        d[1] = v[4] | (v[5] << 8);
        d[2] = 0u;
    }
}

// Sample k (0..5: R, G, B of x, then of x+1) of a packed pair.
template <class T> __device__ __forceinline__ T raw_at(const uint32_t (&d)[3], int k)
{
    if constexpr (sizeof(T) == 2) return (T)((d[k >> 1] >> (16 * (k & 1))) & 0xffffu);
    else return (T)((d[k >> 2] >> (8 * (k & 3))) & 0xffu);
}

// The alpha pair (x, x+1): low byte x, next byte x+1 (0 past the row).
__device__ __forceinline__ uint32_t load_alpha_raw(const uint8_t* p, bool two)
{
    if (two && ((uintptr_t)p & 1) == 0) return (uint32_t)*gmem(reinterpret_cast<const uint16_t*>(p));
    const auto* g = gmem(p);
    return (uint32_t)g[0] | (two ? (uint32_t)g[1] << 8 : 0u);
}

// enfuseMask's exposure + saturation sum of one pixel (zero outside alpha).
template <class T>
__device__ __forceinline__ float weight_of(T p0, T p1, T p2, uint8_t al, const float* __restrict__ lut, double w_exp,
                                          double w_sat)
{
    float wm = 0.f;
    if (!al) return wm;
    if (w_exp > 0.0) wm = lut[gray_code<T>(p0, p1, p2)];
    if (w_sat > 0.0) {
        const T mx = max(p0, max(p1, p2));
        const T mn = min(p0, min(p1, p2));
        float s = 0.f;
        if (mx != mn) {
            const double sum = (double)mx + (double)mn, diff = (double)mx - (double)mn;
            s = (float)(w_sat * (sum <= ImgMax<T>::d ? diff / sum : diff / (2.0 * ImgMax<T>::d - sum)));
        }
        wm = s + wm;
    }
    return wm;
}

template <class T, class M, int NK>
__global__ __launch_bounds__(256) void batch_weights_kernel(const BWeightJob* __restrict__ jobs, const int* __restrict__ tile0,
                                                            int njobs, int total, const BImage* __restrict__ images,
                                                            const float* __restrict__ lut, double w_exp, double w_sat,
                                                            double scale, int rows)
{
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int jn = find_job(tile0, njobs, t);
        const BWeightJob& J = jobs[jn];
        const int local = t - tile0[jn];
        const int x = (local % J.tiles_x) * 128 + 2 * lane;
        const int ybase = (local / J.tiles_x) * rows;
        if (x >= J.w) continue;
        const bool two = x + 1 < J.w;
        const int K = J.K;
        const BImage* im = images + J.img0;
        const int yend = min(J.h, ybase + rows);
        for (int y = ybase + wv; y < yend; y += 4) {
            float w0[NK > 0 ? NK : 1], w1[NK > 0 ? NK : 1];
            float n0 = 0.f, n1 = 0.f;
            uint32_t un = 0;   // union alpha pair
            auto pair_weights = [&](const uint32_t (&d)[3], uint32_t al, float& wa, float& wb) {
                wa = weight_of<T>(raw_at<T>(d, 0), raw_at<T>(d, 1), raw_at<T>(d, 2), (uint8_t)(al & 0xffu), lut, w_exp, w_sat);
                wb = weight_of<T>(raw_at<T>(d, 3), raw_at<T>(d, 4), raw_at<T>(d, 5), (uint8_t)(al >> 8), lut, w_exp, w_sat);
            };
            if constexpr (NK > 0) {
                uint32_t raw[NK][3], ar[NK];
#pragma unroll
                for (int m = 0; m < NK; ++m) {
                    if (m < K) {
                        load_raw<T>(static_cast<const T*>(im[m].img) + (long)y * im[m].ipitch + 3L * x, two, raw[m]);
                        ar[m] = load_alpha_raw(im[m].alpha + (long)y * im[m].apitch + x, two);
                    }
                }
#pragma unroll
                for (int m = 0; m < NK; ++m) {
                    if (m < K) {
                        pair_weights(raw[m], ar[m], w0[m], w1[m]);
                        // the last non-zero alpha of the pixel, as the
                        // reference's union of the alpha masks
                        if (ar[m] & 0xffu) un = (un & 0xff00u) | (ar[m] & 0xffu);
                        if (ar[m] & 0xff00u) un = (un & 0xffu) | (ar[m] & 0xff00u);
                        n0 = w0[m] + n0;
                        n1 = w1[m] + n1;
                    }
                }
            } else {
                for (int m = 0; m < K; ++m) {
                    uint32_t d[3];
                    load_raw<T>(static_cast<const T*>(im[m].img) + (long)y * im[m].ipitch + 3L * x, two, d);
                    const uint32_t al = load_alpha_raw(im[m].alpha + (long)y * im[m].apitch + x, two);
                    float wa, wb;
                    pair_weights(d, al, wa, wb);
                    n0 = wa + n0;
                    n1 = wb + n1;
                    if (al & 0xffu) un = (un & 0xff00u) | (al & 0xffu);
                    if (al & 0xff00u) un = (un & 0xffu) | (al & 0xff00u);
                }
            }
            uint8_t* ur = J.ualpha + (long)y * J.upitch + x;   // x even, rows 64-byte aligned
            if (two) *gmem(reinterpret_cast<uint16_t*>(ur)) = (uint16_t)un;
            else *gmem(ur) = (uint8_t)(un & 0xffu);
            // Normalisation and fixed-point mask level 0 (src/enfuse.h:1529-1546).
            auto store_mask = [&](int m, float wm0, float wm1) {
                const float nm0 = n0 > 0.0 ? (255.0f * wm0) / n0 : J.fallback;
                const float nm1 = n1 > 0.0 ? (255.0f * wm1) / n1 : J.fallback;
                Vec2<M>::store(static_cast<M*>(im[m].mask0) + (long)y * J.mpitch + x, two,
                               from_real<M>((double)nm0 * scale), from_real<M>((double)nm1 * scale));
            };
            if constexpr (NK > 0) {
#pragma unroll
                for (int m = 0; m < NK; ++m)
                    if (m < K) store_mask(m, w0[m], w1[m]);
            } else {
                for (int m = 0; m < K; ++m) {
                    uint32_t d[3];
                    load_raw<T>(static_cast<const T*>(im[m].img) + (long)y * im[m].ipitch + 3L * x, two, d);
                    const uint32_t al = load_alpha_raw(im[m].alpha + (long)y * im[m].apitch + x, two);
                    float wa, wb;
                    pair_weights(d, al, wa, wb);
                    store_mask(m, wa, wb);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Reduce with alpha, no wraparound: the streaming strips of reduce_stream.h
// (src/pyramid.h:166-573), one tile = 4 adjacent strips x a.rows rows.
// ---------------------------------------------------------------------------
// Level-0 source of an image pyramid: the integer image (v << frac) or the
// float image (its log transform).
template <class I, class P, bool WIDE> struct BatchImageSrc { typedef IntImageSrc<I, P, 3, WIDE> type; };
template <bool WIDE> struct BatchImageSrc<float, double, WIDE> { typedef FloatImageSrc<3> type; };

template <class P, class M, class I>
__global__ __launch_bounds__(256) void batch_reduce_kernel(const BReduceJob* __restrict__ jobs, const int* __restrict__ tile0,
                                                           int njobs, int total)
{
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int jn = find_job(tile0, njobs, t);
        const BReduceJob& J = jobs[jn];
        const int local = t - tile0[jn];
        const ReduceArgs a = J.a;
        const int wx = (local % J.groups_x) * 4 + wv;
        const int i0 = (local / J.groups_x) * a.rows;
        if (J.kind == kReduceImage3W) reduce_strip_ring<P, 3, true, typename BatchImageSrc<I, P, true>::type, true>(a, wx, i0);
        else if (J.kind == kReduceImage3) reduce_strip_ring<P, 3, true, typename BatchImageSrc<I, P, false>::type, false>(a, wx, i0);
        else if (J.kind == kReducePlane3) reduce_strip_ring<P, 3, true, PlaneSrc<P, 3>, true>(a, wx, i0);
        else reduce_strip_ring<M, 1, true, PlaneSrc<M, 1>, true>(a, wx, i0);
    }
}

// ---------------------------------------------------------------------------
// Merge + collapse step of one level.
//
// Wave strips: lane L of wave wx owns coarse column j = 62 wx + L - 1 for L
// in [1, 62] and its 2 x 2 fine block; lanes 0 and 63 load the halo columns
// that the owned lanes take through cross-lane shuffles.  Each coarse row
// of the expand is read from the caches once per output row.
// ---------------------------------------------------------------------------

// Polyphase row sums of the expand at coarse column j, no wraparound
// (src/pyramid.h:1199-1411 row forms).
template <class A, class S>
__device__ __forceinline__ void poly_row(S xl, S x, S xr, int j, int cw, S& hA, S& hB)
{
    if (cw == 1) { hA = A::add(A::zero(), A::mul(x, 6)); hB = A::mul(x, 4); }
    else if (j == 0) { hA = A::add(A::add(A::zero(), A::mul(x, 6)), xr); hB = A::mul(A::add(x, xr), 4); }
    else if (j == cw - 1) { hA = A::add(xl, A::mul(x, 6)); hB = A::mul(x, 4); }
    else { hA = A::add(A::add(xl, A::mul(x, 6)), xr); hB = A::mul(A::add(x, xr), 4); }
}

// expand() sums of the 2 x 2 destination block of coarse sample (i, j)
// before the division: o[q][ch], q = dx + 2 dy, from this lane's samples
// c[k][ch] of coarse rows i-1+k (zero outside the level); the horizontal
// neighbours come from the adjacent lanes.
template <class P, int C>
__device__ __forceinline__ void expand_sums(const typename Pix<P>::S (&c)[3][C], int i, int j, int cw, int chh,
                                            typename Pix<P>::S (&o)[4][C])
{
    typedef typename Pix<P>::A A;
    typedef typename Pix<P>::S S;
    S hA[3][C], hB[3][C];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int y = i - 1 + k;
        const bool in = y >= 0 && y < chh;   // uniform over the wave
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            hA[k][ch] = A::zero();
            hB[k][ch] = A::zero();
            if (!in) continue;
            const S x = c[k][ch];
            const S xl = lane_up(x), xr = lane_down(x);
            poly_row<A, S>(xl, x, xr, j, cw, hA[k][ch], hB[k][ch]);
        }
    }
#pragma unroll
    for (int ch = 0; ch < C; ++ch) {
        if (i + 1 < chh) {
            o[0][ch] = A::add(A::add(hA[0][ch], A::mul(hA[1][ch], 6)), hA[2][ch]);
            o[1][ch] = A::add(A::add(hB[0][ch], A::mul(hB[1][ch], 6)), hB[2][ch]);
            o[2][ch] = A::add(hA[1][ch], hA[2][ch]);
            o[3][ch] = A::add(hB[1][ch], hB[2][ch]);
        } else {
            o[0][ch] = A::add(hA[0][ch], A::mul(hA[1][ch], 6));
            o[1][ch] = A::add(hB[0][ch], A::mul(hB[1][ch], 6));
            o[2][ch] = hA[1][ch];
            o[3][ch] = hB[1][ch];
        }
    }
}

// This lane's samples of coarse rows i-1, i, i+1 of a planar level.
template <class P, int C>
__device__ __forceinline__ void load_coarse(const P* p, long pitch, long plane, int i, int j, int chh, bool valid,
                                            typename Pix<P>::S (&c)[3][C])
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int y = i - 1 + k;
#pragma unroll
        for (int ch = 0; ch < C; ++ch)
            c[k][ch] = (valid && y >= 0 && y < chh) ? Pix<P>::load(*gmem(p + ch * plane + (long)y * pitch + j)) : Pix<P>::A::zero();
    }
}

// (double)v / maxpyr for a mask sample v of a normalised fixed-point mask
// pyramid, maxpyr = 255 * 2^mfrac (the ImageMaskMultiplyFunctor coefficient,
// src/enfuse.h:601-620).  For 0 <= v < 2^24 -- every value such a pyramid
// holds -- the quotient by 255 comes from one reciprocal product and one
// fused residual correction, which equals the correctly rounded division on
// that whole range (checked exhaustively by the test suite); the power-of-two
// factor is exact.  Other values take the division.
__device__ __forceinline__ double mask_coeff(int32_t v, double maxpyr, double inv_scale)
{
    if ((uint32_t)v < (1u << 24)) return quot255(v) * inv_scale;
    return (double)v / maxpyr;
}

// SpecializedConvertPyramidToScalarFunctor (src/fixmath.h:236-290).
template <class I>
__device__ __forceinline__ I to_image(int32_t v, int frac, int dither, uint32_t h, int c)
{
    const int32_t q = v >> frac, f = v & ((1 << frac) - 1);
    const int32_t quarter = 1 << (frac - 2), three_quarter = 3 << (frac - 2);
    int32_t r;
    if (f >= quarter && f < three_quarter) {
        if (dither == ENBL_DITHER_DET) r = dither_draw(h, c, frac) <= f ? q + 1 : q;
        else r = f >= (1 << (frac - 1)) ? q + 1 : q;
    } else if (f >= quarter) {
        r = q + 1;
    } else {
        r = q;
    }
    return (I)min(ImgMax<I>::i, max(0, r));
}

// The output sample of a level-0 result value: the dithered integer
// conversion, or the inverse log transform of a float image (src/fixmath.h:
// 320-329).
template <class I, class P>
__device__ __forceinline__ I out_value(P v, int frac, int dither, uint32_t h, int c)
{
    if constexpr (std::is_same<I, float>::value) return (float)log_pyramid_to_float((double)v);
    else return to_image<I>(v, frac, dither, h, c);
}

// Fine rows (X, X+1) of level 0 of a float image's pyramid: the log
// transform of the interleaved samples in double.
struct FloatImageRows {
    static constexpr bool planar = false;
    template <int C, class Args>
    __device__ static __forceinline__ void load(const Args& a, int r, int c, bool both, double (&x0)[C], double (&x1)[C]) {
        const auto* g = gmem(static_cast<const float*>(a.src) + (long)r * a.spitch + (long)c * C);
        float v[2 * C];
#pragma unroll
        for (int k = 0; k < 2 * C; ++k) v[k] = (k < C || both) ? g[k] : 0.f;
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            x0[ch] = float_to_log_pyramid((double)v[ch]);
            x1[ch] = both ? float_to_log_pyramid((double)v[C + ch]) : 0.0;
        }
    }
};

template <class I, class P> struct BatchRows { typedef ImageRows<I, P> type; };
template <> struct BatchRows<float, double> { typedef FloatImageRows type; };

// Fine-level row source of a job entry, in the form the row loaders take.
struct RowSrc {
    const void* src;
    long spitch, splane;
    int frac;
};

// Output of one fine row pair (X, X+1) at level 0: copy under the union alpha
// (src/fixmath.h:2157-2294 with the dithered conversion) and the alpha itself.
template <class P, class I>
__device__ __forceinline__ void emit_pair(const BFuseJob& J, int X, int Y, bool xb, const P (&v0)[3], const P (&v1)[3],
                                          int frac, int dither)
{
    const uint8_t* ur = J.ualpha + (long)Y * J.upitch + X;
    uint8_t u0, u1;
    if (xb) {
        const uint16_t u = *gmem(reinterpret_cast<const uint16_t*>(ur));   // X even, rows 64-byte aligned
        u0 = (uint8_t)(u & 0xffu);
        u1 = (uint8_t)(u >> 8);
    } else {
        u0 = *gmem(ur);
        u1 = 0;
    }
    auto* oa = gmem(J.oalpha + (long)Y * J.oapitch + X);
    oa[0] = u0;
    if (xb) oa[1] = u1;
    I* op = static_cast<I*>(J.out) + (long)Y * J.opitch + 3L * X;
    auto* o = gmem(op);
    I r0[3], r1[3];
    const uint32_t h0 = dither == ENBL_DITHER_DET ? dither_hash(J.seed, X, Y) : 0u;
    const uint32_t h1 = dither == ENBL_DITHER_DET ? dither_hash(J.seed, X + 1, Y) : 0u;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        r0[ch] = out_value<I, P>(v0[ch], frac, dither, h0, ch);
        r1[ch] = out_value<I, P>(v1[ch], frac, dither, h1, ch);
    }
    if constexpr (sizeof(I) == 2) {
        if (xb && u0 && u1 && ((uintptr_t)op & 3) == 0) {
            auto* d = gmem(reinterpret_cast<uint32_t*>(op));
            d[0] = (uint32_t)r0[0] | ((uint32_t)r0[1] << 16);
            d[1] = (uint32_t)r0[2] | ((uint32_t)r1[0] << 16);
            d[2] = (uint32_t)r1[1] | ((uint32_t)r1[2] << 16);
            return;
        }
    }
    if (u0) {
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) o[ch] = r0[ch];
    }
    if (xb && u1) {
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) o[3 + ch] = r1[ch];
    }
}

template <class P, class M, class I, bool LEVEL0>
__global__ __launch_bounds__(256) void batch_fuse_kernel(const BFuseJob* __restrict__ jobs, const int* __restrict__ tile0,
                                                         int njobs, int total, const BPlane* __restrict__ planes,
                                                         double maxpyr, double inv_scale, int frac, int dither, int rows)
{
    typedef Pix<P> Px;
    typedef typename Px::A A;
    typedef typename Px::S S;
    typedef typename std::conditional<LEVEL0, typename BatchRows<I, P>::type, PlaneRows<P>>::type FINE;
    typedef typename std::conditional<LEVEL0, ImageLoad<I, P>, PlaneLoad<P>>::type FINE1;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int jn = find_job(tile0, njobs, t);
        const BFuseJob& J = jobs[jn];
        const int local = t - tile0[jn];
        const int fw = J.fw, fh = J.fh, cw = J.cw, chh = J.ch;
        const BPlane* pl = planes + J.plane0;   // per image: fine, coarse, mask
        if (cw == 0) {
            // Coarsest level of the stack: the Laplacian level is the Gaussian
            // level and nothing is added.  One fine sample per thread.
            const int j = (local % J.tiles_x) * 64 + lane;
            const int ybase = (local / J.tiles_x) * rows;
            if (j >= fw) continue;
            const int yend = min(fh, ybase + rows);
            for (int y = ybase + wv; y < yend; y += 4) {
                S acc[3];
                for (int m = 0; m < J.K; ++m) {
                    const BPlane mk = pl[3 * m + 2];
                    const double coeff = mask_coeff(*gmem(static_cast<const M*>(mk.p) + (long)y * mk.pitch + j), maxpyr, inv_scale);
#pragma unroll
                    for (int ch = 0; ch < 3; ++ch) {
                        const S v = Px::load(from_real<P>((double)FINE1::get(pl[3 * m], y, j, ch, frac) * coeff));
                        acc[ch] = m == 0 ? v : Px::load(Px::wrap(A::add(v, acc[ch])));
                    }
                }
                if (LEVEL0) {
                    const uint8_t ua = *gmem(J.ualpha + (long)y * J.upitch + j);
                    *gmem(J.oalpha + (long)y * J.oapitch + j) = ua;
                    if (ua) {
                        auto* o = gmem(static_cast<I*>(J.out) + (long)y * J.opitch + 3L * j);
                        const uint32_t hh = dither == ENBL_DITHER_DET ? dither_hash(J.seed, j, y) : 0u;
#pragma unroll
                        for (int ch = 0; ch < 3; ++ch) o[ch] = out_value<I, P>((P)acc[ch], frac, dither, hh, ch);
                    }
                } else {
#pragma unroll
                    for (int ch = 0; ch < 3; ++ch) *gmem(static_cast<P*>(J.res) + ch * J.rplane + (long)y * J.rpitch + j) = (P)acc[ch];
                }
            }
            continue;
        }
        const int wx = (local % J.tiles_x) * 4 + wv;
        if (wx * kOwned >= cw) continue;   // the whole wave is beyond the level
        const int i0 = (local / J.tiles_x) * rows;
        const int i1 = min(chh, i0 + rows);
        const int j = wx * kOwned + lane - 1;
        const bool valid = j >= 0 && j < cw;
        const bool owned = valid && lane >= 1 && lane <= kOwned;
        const int X = 2 * j;
        const bool xb = X + 1 < fw;
        const int hdA = expand_hsum(X, cw, false, true), hdB = expand_hsum(X + 1, cw, false, true);
        for (int i = i0; i < i1; ++i) {
            const int Y = 2 * i;
            const bool yb = Y + 1 < fh;
            const int v0 = expand_vsum(Y, chh), v1 = expand_vsum(Y + 1, chh);
            const int dv[4] = {hdA * v0, hdB * v0, hdA * v1, hdB * v1};
            // The inputs of image m at this coarse row.
            struct In {
                S c[3][3];     // coarse rows i-1..i+1, own column
                S f[4][3];     // fine samples of the 2 x 2 block
                int32_t mv[4]; // mask samples of the 2 x 2 block
            };
            auto load_in = [&](int m, In& D) {
                const BPlane fs = pl[3 * m];
                const BPlane cs = pl[3 * m + 1];
                const BPlane mk = pl[3 * m + 2];
                load_coarse<P, 3>(static_cast<const P*>(cs.p), cs.pitch, cs.plane, i, j, chh, valid, D.c);
                const RowSrc rs{fs.p, fs.pitch, fs.plane, frac};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    D.mv[q] = 0;
#pragma unroll
                    for (int ch = 0; ch < 3; ++ch) D.f[q][ch] = A::zero();
                }
                if (valid) {
                    const M* mrow = static_cast<const M*>(mk.p) + (long)Y * mk.pitch + X;
                    FINE::template load<3>(rs, Y, X, xb, D.f[0], D.f[1]);
                    Vec2<M>::load(mrow, xb, D.mv[0], D.mv[1]);
                    if (yb) {
                        FINE::template load<3>(rs, Y + 1, X, xb, D.f[2], D.f[3]);
                        Vec2<M>::load(mrow + mk.pitch, xb, D.mv[2], D.mv[3]);
                    }
                }
            };
            S acc[4][3];
            for (int m = 0; m < J.K; ++m) {
                In cur;
                load_in(m, cur);
                S o[4][3];
                expand_sums<P, 3>(cur.c, i, j, cw, chh, o);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    // Laplacian value (src/pyramid.h:1705-1707), mask multiply
                    // (src/enfuse.h:601-620) and accumulation (:1576-1588).
                    const double coeff = mask_coeff(cur.mv[q], maxpyr, inv_scale);
#pragma unroll
                    for (int ch = 0; ch < 3; ++ch) {
                        const P lapv = Px::wrap(A::sub(cur.f[q][ch], A::div(o[q][ch], dv[q])));
                        const S v = Px::load(from_real<P>((double)lapv * coeff));
                        acc[q][ch] = m == 0 ? v : Px::load(Px::wrap(A::add(v, acc[q][ch])));
                    }
                }
            }
            // Collapse step: + expand(result[l+1]), saturating (src/pyramid.h:1837-1886).
            S rcv[3][3], o[4][3];
            load_coarse<P, 3>(static_cast<const P*>(J.rcoarse), J.rcpitch, J.rcplane, i, j, chh, valid, rcv);
            expand_sums<P, 3>(rcv, i, j, cw, chh, o);
            P outv[4][3];
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) outv[q][ch] = Px::sat(A::add(acc[q][ch], A::div(o[q][ch], dv[q])));
            if (!owned) continue;
            if (LEVEL0) {
                emit_pair<P, I>(J, X, Y, xb, outv[0], outv[1], frac, dither);
                if (yb) emit_pair<P, I>(J, X, Y + 1, xb, outv[2], outv[3], frac, dither);
            } else {
                P* r0 = static_cast<P*>(J.res) + (long)Y * J.rpitch + X;
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) {
                    Vec2<P>::store(r0 + ch * J.rplane, xb, outv[0][ch], outv[1][ch]);
                    if (yb) Vec2<P>::store(r0 + ch * J.rplane + J.rpitch, xb, outv[2][ch], outv[3][ch]);
                }
            }
        }
    }
}

// Fine samples (X, X+1) of row Y of level 0 of an image pyramid, formed from
// the interleaved image: (P)v << frac for integer images, the log transform
// of src/fixmath.h:165-186 in double for float images.
template <class I, class P> struct L0Fine {
    typedef typename Pix<P>::S S;
    __device__ static __forceinline__ void load(const void* img, long ip, int Y, int X, bool xb, int frac, S (&x0)[3],
                                                S (&x1)[3]) {
        const RowSrc rs{img, ip, 0, frac};
        BatchRows<I, P>::type::template load<3>(rs, Y, X, xb, x0, x1);
    }
};

// Output of one fine pixel pair (X, X+1) of row Y: the union alpha, and the
// image under it, SpecializedConvertPyramidToScalarFunctor with the dither
// for integer images (src/fixmath.h:236-290), the inverse log transform for
// float images (:320-329).
template <class P, class I>
__device__ __forceinline__ void emit_l0(const L0FuseArgs& a, int X, int Y, bool xb, const P (&v0)[3], const P (&v1)[3])
{
    const uint8_t* ur = a.ualpha + (long)Y * a.upitch + X;
    const uint8_t u0 = *gmem(ur), u1 = xb ? *gmem(ur + 1) : (uint8_t)0;
    auto* oa = gmem(a.oalpha + (long)Y * a.oapitch + X);
    oa[0] = u0;
    if (xb) oa[1] = u1;
    auto* o = gmem(static_cast<I*>(a.out) + (long)Y * a.opitch + 3L * X);
    if constexpr (std::is_same<I, float>::value) {
        if (u0) {
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) o[ch] = (float)log_pyramid_to_float(v0[ch]);
        }
        if (u1) {
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) o[3 + ch] = (float)log_pyramid_to_float(v1[ch]);
        }
    } else {
        const uint32_t h0 = a.dither == ENBL_DITHER_DET ? dither_hash(a.seed, X, Y) : 0u;
        const uint32_t h1 = a.dither == ENBL_DITHER_DET ? dither_hash(a.seed, X + 1, Y) : 0u;
        if (u0) {
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) o[ch] = to_image<I>(v0[ch], a.frac, a.dither, h0, ch);
        }
        if (u1) {
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) o[3 + ch] = to_image<I>(v1[ch], a.frac, a.dither, h1, ch);
        }
    }
}

// One coarse row i of one image of the level-0 merge (see
// enfuse_level0_kernel): the Laplacian values of the 2 x 2 fine block, times
// the image's mask values, added to acc in image order.  c: the image's
// Gaussian level-1 rows i-1, i, i+1.  P2: every divisor of the wave's expand
// is a power of two (interior rows and columns), so the divisions are exact
// products with the reciprocals.
template <class P, class M, class I, bool G0, bool P2>
__device__ __forceinline__ void level0_image_row(const L0FuseArgs& a, double inv_scale, int m, int i, int j, bool valid,
                                                 int X, bool xb, const int (&dv)[4], const float (&nv)[4],
                                                 const typename Pix<P>::S (&c)[3][3], typename Pix<P>::S (&acc)[4][3])
{
    typedef Pix<P> Px;
    typedef typename Px::A A;
    typedef typename Px::S S;
    const int Y = 2 * i;
    const bool yb = Y + 1 < a.fh;
    auto qdiv = [&](S v, int q) -> S {
        if constexpr (P2 && std::is_same<P, double>::value) {
            return v * __builtin_amdgcn_ldexp(1.0, -(int)__builtin_ctz((unsigned)dv[q]));
        } else if constexpr (P2) {
            const int sh = __builtin_ctz((unsigned)dv[q]);
            return (S)((v + ((v >> 31) & ((1 << sh) - 1))) >> sh);
        } else {
            return A::div(v, dv[q]);
        }
    };
    S f[4][3];
    int32_t mv[4] = {0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) f[q][ch] = A::zero();
    if (valid) {
        const float* wrow = a.wt[m] + (long)Y * a.wpitch + X;
        float wv[4];
        wv[0] = *gmem(wrow);
        wv[1] = xb ? *gmem(wrow + 1) : 0.f;
        wv[2] = yb ? *gmem(wrow + a.wpitch) : 0.f;
        wv[3] = xb && yb ? *gmem(wrow + a.wpitch + 1) : 0.f;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            mv[q] = (int32_t)from_real<M>(norm_mask_value(wv[q], nv[q], a.fallback, a.hard, a.mscale));
        if constexpr (G0) {
            const RowSrc rs{a.g0[m], a.g0pitch, a.g0plane, 0};
            PlaneRows<P>::template load<3>(rs, Y, X, xb, f[0], f[1]);
            if (yb) PlaneRows<P>::template load<3>(rs, Y + 1, X, xb, f[2], f[3]);
        } else {
            L0Fine<I, P>::load(a.img[m], a.ipitch[m], Y, X, xb, a.frac, f[0], f[1]);
            if (yb) L0Fine<I, P>::load(a.img[m], a.ipitch[m], Y + 1, X, xb, a.frac, f[2], f[3]);
        }
    }
    S o[4][3];
    expand_sums<P, 3>(c, i, j, a.cw, a.ch, o);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        // Laplacian value (src/pyramid.h:1705-1707), mask multiply
        // (src/enfuse.h:601-620) and accumulation (:1576-1588).
        const double coeff = mask_coeff(mv[q], a.maxpyr, inv_scale);
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const P lapv = Px::wrap(A::sub(f[q][ch], qdiv(o[q][ch], q)));
            const S v = Px::load(from_real<P>((double)lapv * coeff));
            acc[q][ch] = m == 0 ? v : Px::load(Px::wrap(A::add(v, acc[q][ch])));
        }
    }
}

// Collapse step of coarse row i: + expand(result[1]), saturating
// (src/pyramid.h:1837-1886), and the output of the 2 x 2 block.
template <class P, class I, bool P2>
__device__ __forceinline__ void level0_collapse_row(const L0FuseArgs& a, int i, int j, bool owned, int X, bool xb,
                                                    const int (&dv)[4], const typename Pix<P>::S (&rcv)[3][3],
                                                    const typename Pix<P>::S (&acc)[4][3])
{
    typedef Pix<P> Px;
    typedef typename Px::A A;
    typedef typename Px::S S;
    const int Y = 2 * i;
    const bool yb = Y + 1 < a.fh;
    auto qdiv = [&](S v, int q) -> S {
        if constexpr (P2 && std::is_same<P, double>::value) {
            return v * __builtin_amdgcn_ldexp(1.0, -(int)__builtin_ctz((unsigned)dv[q]));
        } else if constexpr (P2) {
            const int sh = __builtin_ctz((unsigned)dv[q]);
            return (S)((v + ((v >> 31) & ((1 << sh) - 1))) >> sh);
        } else {
            return A::div(v, dv[q]);
        }
    };
    S o[4][3];
    expand_sums<P, 3>(rcv, i, j, a.cw, a.ch, o);
    P outv[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) outv[q][ch] = Px::sat(A::add(acc[q][ch], qdiv(o[q][ch], q)));
    if (!owned) return;
    emit_l0<P, I>(a, X, Y, xb, outv[0], outv[1]);
    if (yb) emit_l0<P, I>(a, X, Y + 1, xb, outv[2], outv[3]);
}

// One coarse row y of a planar level into c (zero outside the level and on
// lanes without a column).
template <class P>
__device__ __forceinline__ void load_coarse_row(const P* p, long pitch, long plane, int y, int j, int chh, bool valid,
                                                typename Pix<P>::S (&c)[3])
{
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
        c[ch] = (valid && y >= 0 && y < chh) ? Pix<P>::load(*gmem(p + ch * plane + (long)y * pitch + j)) : Pix<P>::A::zero();
}

// Wave strips as in batch_fuse_kernel: lane L of wave wx owns coarse column
// j = 62 wx + L - 1 (L in [1, 62]) and its 2 x 2 fine block; a block is four
// waves side by side over R coarse rows.  Each wave walks the images in
// order, and for each image its R rows with a rolling window of three
// Gaussian level-1 rows, so a level-1 row is read once per image and strip
// (plus the two halo rows) while the R rows' sums stay in registers; then the
// collapse of the R rows, again with a rolling window of result rows.
// G0: level 0 read from the converted planes a.g0 instead of the images.
template <class P, class M, class I, bool G0, int R>
__global__ __launch_bounds__(256) void enfuse_level0_kernel(L0FuseArgs a, double inv_scale)
{
    typedef typename Pix<P>::S S;
    const int lane = threadIdx.x & 63;
    const int wx = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int cw = a.cw, chh = a.ch;
    if (wx * kOwned >= cw) return;   // the whole wave is beyond the level
    const int i0 = blockIdx.y * R;
    const int nr = min(R, chh - i0);
    const int j = wx * kOwned + lane - 1;
    const bool valid = j >= 0 && j < cw;
    const bool owned = valid && lane >= 1 && lane <= kOwned;
    const int X = 2 * j;
    const bool xb = X + 1 < a.fw;
    const int hdA = expand_hsum(X, cw, false, true), hdB = expand_hsum(X + 1, cw, false, true);
    int dv[R][4];
    bool p2[R];
    float nv[R][4];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = i0 + r, Y = 2 * i;
        const int v0 = expand_vsum(Y, chh), v1 = expand_vsum(Y + 1, chh);
        dv[r][0] = hdA * v0; dv[r][1] = hdB * v0; dv[r][2] = hdA * v1; dv[r][3] = hdB * v1;
        // (only the owned lanes' results are kept)
        bool q2 = true;
#pragma unroll
        for (int q = 0; q < 4; ++q) q2 = q2 && dv[r][q] > 0 && (dv[r][q] & (dv[r][q] - 1)) == 0;
        p2[r] = wave_all(q2 || !owned || r >= nr);
        // the norm of the 2 x 2 block, shared by the images' mask values
        const bool yb = Y + 1 < a.fh;
#pragma unroll
        for (int q = 0; q < 4; ++q) nv[r][q] = 0.f;
        if (valid && r < nr) {
            const float* nrow = a.norm + (long)Y * a.npitch + X;
            nv[r][0] = *gmem(nrow);
            nv[r][1] = xb ? *gmem(nrow + 1) : 0.f;
            if (yb) {
                nv[r][2] = *gmem(nrow + a.npitch);
                nv[r][3] = xb ? *gmem(nrow + a.npitch + 1) : 0.f;
            }
        }
    }
    S acc[R][4][3];
    for (int m = 0; m < a.K; ++m) {
        const P* g1 = static_cast<const P*>(a.g1[m]);
        S c[3][3];
        load_coarse_row<P>(g1, a.gpitch, a.gplane, i0 - 1, j, chh, valid, c[0]);
        load_coarse_row<P>(g1, a.gpitch, a.gplane, i0, j, chh, valid, c[1]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r >= nr) break;
            const int i = i0 + r;
            load_coarse_row<P>(g1, a.gpitch, a.gplane, i + 1, j, chh, valid, c[2]);
            if (p2[r]) level0_image_row<P, M, I, G0, true>(a, inv_scale, m, i, j, valid, X, xb, dv[r], nv[r], c, acc[r]);
            else level0_image_row<P, M, I, G0, false>(a, inv_scale, m, i, j, valid, X, xb, dv[r], nv[r], c, acc[r]);
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                c[0][ch] = c[1][ch];
                c[1][ch] = c[2][ch];
            }
        }
    }
    const P* rc = static_cast<const P*>(a.rc);
    S rcv[3][3];
    load_coarse_row<P>(rc, a.rcpitch, a.rcplane, i0 - 1, j, chh, valid, rcv[0]);
    load_coarse_row<P>(rc, a.rcpitch, a.rcplane, i0, j, chh, valid, rcv[1]);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (r >= nr) break;
        const int i = i0 + r;
        load_coarse_row<P>(rc, a.rcpitch, a.rcplane, i + 1, j, chh, valid, rcv[2]);
        if (p2[r]) level0_collapse_row<P, I, true>(a, i, j, owned, X, xb, dv[r], rcv, acc[r]);
        else level0_collapse_row<P, I, false>(a, i, j, owned, X, xb, dv[r], rcv, acc[r]);
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            rcv[0][ch] = rcv[1][ch];
            rcv[1][ch] = rcv[2][ch];
        }
    }
}

}  // namespace

int launch_enfuse_level0(int itype, L0FuseArgs a, hipStream_t s)
{
    if (a.K < 1 || a.K > kL0Max || a.cw < 1 || a.ch < 1) return ENBL_E_ARG;
    constexpr int R = 3;   // coarse rows per wave (their sums stay in registers)
    const int waves_x = (a.cw + kOwned - 1) / kOwned;
    const dim3 grid((waves_x + 3) / 4, (a.ch + R - 1) / R);
    const double inv_scale = 255.0 / a.maxpyr;   // 2^-mfrac, exact
    // (8- and 16-bit stacks keep their stored level 0 in the per-stack path;
    // the integer instances would form it as (P)v << frac)
    if (itype == ENBL_F32 && a.g0[0])
        hipLaunchKernelGGL((enfuse_level0_kernel<double, int32_t, float, true, R>), grid, dim3(256), 0, s, a, inv_scale);
    else if (itype == ENBL_F32)
        hipLaunchKernelGGL((enfuse_level0_kernel<double, int32_t, float, false, R>), grid, dim3(256), 0, s, a, inv_scale);
    else return ENBL_E_UNSUPPORTED;
    return 0;
}

int launch_batch_weights(int itype, int max_k, const BWeightJob* jobs, const int* tile0, int njobs, int total,
                         const BImage* images, const float* lut, double w_exposure, double w_saturation, double scale, int rows, int grid,
                         hipStream_t s)
{
    if (total <= 0) return 0;
#define ENBL_BW(T, M, NK) hipLaunchKernelGGL((batch_weights_kernel<T, M, NK>), dim3(grid), dim3(256), 0, s, jobs, tile0, njobs, \
                                             total, images, lut, w_exposure, w_saturation, scale, rows)
#define ENBL_BW_K(T, M)                                  \
    do {                                                 \
        if (max_k <= 4) ENBL_BW(T, M, 4);                \
        else if (max_k <= 9) ENBL_BW(T, M, 9);           \
        else if (max_k <= 12) ENBL_BW(T, M, 12);         \
        else if (max_k <= 16) ENBL_BW(T, M, 16);         \
        else ENBL_BW(T, M, 0);                           \
    } while (0)
    if (itype == ENBL_U8) ENBL_BW_K(uint8_t, int16_t);
    else if (itype == ENBL_U16) ENBL_BW_K(uint16_t, int32_t);
    else return ENBL_E_UNSUPPORTED;
#undef ENBL_BW_K
#undef ENBL_BW
    return 0;
}

int launch_batch_reduce(int itype, const BReduceJob* jobs, const int* tile0, int njobs, int total, int grid, hipStream_t s)
{
    if (total <= 0) return 0;
    if (itype == ENBL_U8)
        hipLaunchKernelGGL((batch_reduce_kernel<int16_t, int16_t, uint8_t>), dim3(grid), dim3(256), 0, s, jobs, tile0, njobs, total);
    else if (itype == ENBL_U16)
        hipLaunchKernelGGL((batch_reduce_kernel<int32_t, int32_t, uint16_t>), dim3(grid), dim3(256), 0, s, jobs, tile0, njobs, total);
    else if (itype == ENBL_F32)
        hipLaunchKernelGGL((batch_reduce_kernel<double, int32_t, float>), dim3(grid), dim3(256), 0, s, jobs, tile0, njobs, total);
    else return ENBL_E_UNSUPPORTED;
    return 0;
}

int launch_batch_fuse(int itype, int frac, int level0, const BFuseJob* jobs, const int* tile0, int njobs, int total,
                      const BPlane* planes, double maxpyr, int dither, int rows, int grid, hipStream_t s)
{
    if (total <= 0) return 0;
    const double inv_scale = 255.0 / maxpyr;   // 2^-mfrac, exact
#define ENBL_BFUSE(P, M, I)                                                                                          \
    do {                                                                                                             \
        if (level0) hipLaunchKernelGGL((batch_fuse_kernel<P, M, I, true>), dim3(grid), dim3(256), 0, s, jobs, tile0, njobs, total, \
                                       planes, maxpyr, inv_scale, frac, dither, rows);                                       \
        else hipLaunchKernelGGL((batch_fuse_kernel<P, M, I, false>), dim3(grid), dim3(256), 0, s, jobs, tile0, njobs, total,   \
                                planes, maxpyr, inv_scale, frac, dither, rows);                                              \
    } while (0)
    if (itype == ENBL_U8) ENBL_BFUSE(int16_t, int16_t, uint8_t);
    else if (itype == ENBL_U16) ENBL_BFUSE(int32_t, int32_t, uint16_t);
    else if (itype == ENBL_F32) ENBL_BFUSE(double, int32_t, float);
    else return ENBL_E_UNSUPPORTED;
#undef ENBL_BFUSE
    return 0;
}

}  // namespace enbl
