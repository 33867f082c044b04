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
        return Pix<P>::load(static_cast<const P*>(b.p)[ch * b.plane + (long)y * b.pitch + x]);
    }
};
template <class I, class P> struct ImageLoad {
    typedef typename Pix<P>::S S;
    __device__ static __forceinline__ S get(const BPlane& b, int y, int x, int ch, int frac) {
        return (S)(P)((P)static_cast<const I*>(b.p)[(long)y * b.pitch + 3L * x + ch] << frac);
    }
};

// "average" grayscale projector (src/mga.h:188-193) with the integer
// fromRealPromote of the image type.
template <class T> __device__ __forceinline__ int gray_code(T p0, T p1, T p2)
{
    const double v = ((double)p0 + (double)p1 + (double)p2) / 3.0;
    if (v < 0.0) return 0;
    return v > ImgMax<T>::d ? ImgMax<T>::i : (int)(int32_t)(v + 0.5);
}

// ---------------------------------------------------------------------------
// Weights: exposure (host table by gray code) then saturation, float sums in
// the reference's order (src/enfuse.h:1000-1008, :1131-1135); norm += w in
// image order (:1377-1385); normalisation and fixed-point mask level 0
// (:1529-1546, src/fixmath.h:188-194).
// ---------------------------------------------------------------------------
constexpr int kRegK = 16;   // weights of up to this many images stay in registers

template <class T>
__device__ __forceinline__ float pixel_weight(const BWeightJob& J, int m, int x, int y, const float* __restrict__ lut,
                                              double w_exp, double w_sat, uint8_t& al)
{
    al = J.alphas[m][(long)y * J.apitch[m] + x];
    float wm = 0.f;
    if (!al) return wm;
    const T* p = static_cast<const T*>(J.imgs[m]) + (long)y * J.ipitch[m] + 3L * x;
    const T p0 = p[0], p1 = p[1], p2 = p[2];
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

template <class T, class M>
__global__ __launch_bounds__(256) void batch_weights_kernel(const BWeightJob* __restrict__ jobs, const int* __restrict__ tile0,
                                                            int njobs, int total, const float* __restrict__ lut,
                                                            double w_exp, double w_sat, double scale, int rows)
{
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int jn = find_job(tile0, njobs, t);
        const BWeightJob& J = jobs[jn];
        const int local = t - tile0[jn];
        const int x = (local % J.tiles_x) * 64 + lx;
        const int ybase = (local / J.tiles_x) * rows;
        if (x >= J.w) continue;
        const int yend = min(J.h, ybase + rows);
        for (int y = ybase + ly; y < yend; y += 4) {
            float wreg[kRegK];
            float norm = 0.f;
            uint8_t ua = 0;
#pragma unroll
            for (int m = 0; m < kRegK; ++m) {
                if (m >= J.K) break;
                uint8_t al;
                wreg[m] = pixel_weight<T>(J, m, x, y, lut, w_exp, w_sat, al);
                if (al) ua = al;
                norm = wreg[m] + norm;
            }
            for (int m = kRegK; m < J.K; ++m) {
                uint8_t al;
                const float wm = pixel_weight<T>(J, m, x, y, lut, w_exp, w_sat, al);
                if (al) ua = al;
                norm = wm + norm;
            }
            J.ualpha[(long)y * J.upitch + x] = ua;
#pragma unroll
            for (int m = 0; m < kRegK; ++m) {
                if (m >= J.K) break;
                const float nm = norm > 0.0 ? (255.0f * wreg[m]) / norm : J.fallback;
                static_cast<M*>(J.mask0[m])[(long)y * J.mpitch + x] = from_real<M>((double)nm * scale);
            }
            for (int m = kRegK; m < J.K; ++m) {
                uint8_t al;
                const float wm = pixel_weight<T>(J, m, x, y, lut, w_exp, w_sat, al);
                const float nm = norm > 0.0 ? (255.0f * wm) / norm : J.fallback;
                static_cast<M*>(J.mask0[m])[(long)y * J.mpitch + x] = from_real<M>((double)nm * scale);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Reduce with alpha, no wraparound (src/pyramid.h:161-560 with a mask):
// zero-extended 1-4-6-4-1 sums normalised by the in-mask weight, the
// reference's boundary forms and operation order; one output per thread.
// ---------------------------------------------------------------------------
template <class P, int C, class LOAD>
__device__ __forceinline__ void reduce_rows(const BReduceJob& J, int j, int ybase, int rows, int ly, int frac)
{
    typedef Pix<P> Px;
    typedef typename Px::A A;
    typedef typename Px::S S;
    const int sw = J.sw, sh = J.sh;
    const int c = 2 * j;
    const int rcase = c == sw - 1 ? 1 : (c == sw - 2 ? 2 : 0);
    int col[5];
    bool in[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int x = c - 2 + k;
        in[k] = x >= 0 && x < sw;
        col[k] = x < 0 ? 0 : (x >= sw ? sw - 1 : x);
    }
    P* dst = static_cast<P*>(J.dst);
    const int iend = min(J.dh, ybase + rows);
    for (int i = ybase + ly; i < iend; i += 4) {
        const int r = 2 * i;
        int rowi[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const int y = r - 2 + k;
            rowi[k] = y < 0 ? 0 : (y >= sh ? sh - 1 : y);
        }
        // All window loads first (clamped indices stay inside the level).
        uint8_t al[5][5];
        S raw[5][5][C];
#pragma unroll
        for (int k = 0; k < 5; ++k)
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                al[k][q] = J.salpha[(long)rowi[k] * J.sapitch + col[q]];
#pragma unroll
                for (int ch = 0; ch < C; ++ch) raw[k][q][ch] = LOAD::get(J.src, rowi[k], col[q], ch, frac);
            }
        S hrow[5][C];
        int mrow[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const int y = r - 2 + k;
            mrow[k] = 0;
#pragma unroll
            for (int ch = 0; ch < C; ++ch) hrow[k][ch] = A::zero();
            if (y < 0 || y >= sh) continue;
            int m[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) m[q] = (in[q] && al[k][q]) ? 1 : 0;
            int hm = m[0] + 4 * m[1] + 6 * m[2];
            if (rcase != 1) hm += 4 * m[3];
            if (rcase == 0) hm += m[4];
            mrow[k] = hm;
#pragma unroll
            for (int ch = 0; ch < C; ++ch) {
                S v[5];
#pragma unroll
                for (int q = 0; q < 5; ++q) v[q] = m[q] ? raw[k][q][ch] : A::zero();
                const S a0 = A::add(v[0], A::mul(v[1], 4));
                if (rcase == 1) hrow[k][ch] = A::add(a0, A::mul(v[2], 6));
                else if (rcase == 2) hrow[k][ch] = A::add(A::add(a0, A::mul(v[2], 6)), A::mul(v[3], 4));
                else hrow[k][ch] = A::add(A::add(A::add(a0, A::mul(v[2], 6)), A::mul(v[3], 4)), v[4]);
            }
        }
        const int q0 = r == 0 ? 0 : mrow[0] + 4 * mrow[1];
        int ap;
        if (r == sh - 1) ap = q0 + 6 * mrow[2];
        else if (r == sh - 2) ap = q0 + 6 * mrow[2] + 4 * mrow[3];
        else ap = q0 + 6 * mrow[2] + 4 * mrow[3] + mrow[4];
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            const S a0 = r == 0 ? A::zero() : A::add(hrow[0][ch], A::mul(hrow[1][ch], 4));
            S v;
            if (r == sh - 1) v = A::add(a0, A::mul(hrow[2][ch], 6));
            else if (r == sh - 2) v = A::add(A::add(a0, A::mul(hrow[2][ch], 6)), A::mul(hrow[3][ch], 4));
            else v = A::add(A::add(A::add(a0, A::mul(hrow[2][ch], 6)), A::mul(hrow[3][ch], 4)), hrow[4][ch]);
            dst[ch * J.dplane + (long)i * J.dpitch + j] = ap ? Px::wrap(A::div(v, ap)) : (P)0;
        }
        if (J.dalpha) J.dalpha[(long)i * J.dapitch + j] = ap ? 255 : 0;
    }
}

template <class P, class I>
__global__ __launch_bounds__(256) void batch_reduce_kernel(const BReduceJob* __restrict__ jobs, const int* __restrict__ tile0,
                                                           int njobs, int total, int rows, int frac)
{
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int jn = find_job(tile0, njobs, t);
        const BReduceJob& J = jobs[jn];
        const int local = t - tile0[jn];
        const int j = (local % J.tiles_x) * 64 + lx;
        const int ybase = (local / J.tiles_x) * rows;
        if (j >= J.dw) continue;
        if (J.channels == 1) reduce_rows<P, 1, PlaneLoad<P>>(J, j, ybase, rows, ly, frac);
        else if (J.src_image) reduce_rows<P, 3, ImageLoad<I, P>>(J, j, ybase, rows, ly, frac);
        else reduce_rows<P, 3, PlaneLoad<P>>(J, j, ybase, rows, ly, frac);
    }
}

// ---------------------------------------------------------------------------
// Merge + collapse step of one level.
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

// expand() of one coarse sample's 2x2 destination block before the division:
// o[q], q = dx + 2 dy, from the 3x3 coarse neighbourhood.
template <class P, class GET>
__device__ __forceinline__ void expand_block(GET&& get, int i, int j, int cw, int chh, typename Pix<P>::S o[4])
{
    typedef typename Pix<P>::A A;
    typedef typename Pix<P>::S S;
    S hA[3], hB[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        hA[k] = A::zero();
        hB[k] = A::zero();
        const int y = i - 1 + k;
        if (y < 0 || y >= chh) continue;
        const S x = get(y, j);
        const S xl = j > 0 ? get(y, j - 1) : A::zero();
        const S xr = j + 1 < cw ? get(y, j + 1) : A::zero();
        poly_row<A, S>(xl, x, xr, j, cw, hA[k], hB[k]);
    }
    if (i + 1 < chh) {
        o[0] = A::add(A::add(hA[0], A::mul(hA[1], 6)), hA[2]);
        o[1] = A::add(A::add(hB[0], A::mul(hB[1], 6)), hB[2]);
        o[2] = A::add(hA[1], hA[2]);
        o[3] = A::add(hB[1], hB[2]);
    } else {
        o[0] = A::add(hA[0], A::mul(hA[1], 6));
        o[1] = A::add(hB[0], A::mul(hB[1], 6));
        o[2] = hA[1];
        o[3] = hB[1];
    }
}

// SpecializedConvertPyramidToScalarFunctor (src/fixmath.h:236-290).
template <class I>
__device__ __forceinline__ I to_image(int32_t v, int frac, int dither, uint64_t seed, int x, int y, int c)
{
    const int32_t q = v >> frac, f = v & ((1 << frac) - 1);
    const int32_t quarter = 1 << (frac - 2), three_quarter = 3 << (frac - 2);
    int32_t r;
    if (f >= quarter && f < three_quarter) {
        if (dither == ENBL_DITHER_DET) r = dither_rand(seed, x, y, c, frac) <= f ? q + 1 : q;
        else r = f >= (1 << (frac - 1)) ? q + 1 : q;
    } else if (f >= quarter) {
        r = q + 1;
    } else {
        r = q;
    }
    return (I)min(ImgMax<I>::i, max(0, r));
}

template <class P, class M, class I, bool LEVEL0>
__global__ __launch_bounds__(256) void batch_fuse_kernel(const BFuseJob* __restrict__ jobs, const int* __restrict__ tile0,
                                                         int njobs, int total, double maxpyr, int frac, int dither,
                                                         int rows)
{
    typedef Pix<P> Px;
    typedef typename Px::A A;
    typedef typename Px::S S;
    typedef typename std::conditional<LEVEL0, ImageLoad<I, P>, PlaneLoad<P>>::type FINE;
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int jn = find_job(tile0, njobs, t);
        const BFuseJob& J = jobs[jn];
        const int local = t - tile0[jn];
        const int j = (local % J.tiles_x) * 64 + lx;
        const int ybase = (local / J.tiles_x) * rows;
        const int fw = J.fw, fh = J.fh, cw = J.cw, chh = J.ch;
        if (cw == 0) {
            // Coarsest level of the stack: the Laplacian level is the Gaussian
            // level and nothing is added.  One fine sample per thread.
            if (j >= fw) continue;
            const int yend = min(fh, ybase + rows);
            for (int y = ybase + ly; y < yend; y += 4) {
                S acc[3];
                for (int m = 0; m < J.K; ++m) {
                    const BPlane mk = J.mask[m];
                    const double coeff = (double)static_cast<const M*>(mk.p)[(long)y * mk.pitch + j] / maxpyr;
#pragma unroll
                    for (int ch = 0; ch < 3; ++ch) {
                        const S v = Px::load(from_real<P>((double)FINE::get(J.fine[m], y, j, ch, frac) * coeff));
                        acc[ch] = m == 0 ? v : Px::load(Px::wrap(A::add(v, acc[ch])));
                    }
                }
                if (LEVEL0) {
                    const uint8_t ua = J.ualpha[(long)y * J.upitch + j];
                    J.oalpha[(long)y * J.oapitch + j] = ua;
                    if (ua) {
                        I* o = static_cast<I*>(J.out) + (long)y * J.opitch + 3L * j;
#pragma unroll
                        for (int ch = 0; ch < 3; ++ch) o[ch] = to_image<I>(acc[ch], frac, dither, J.seed, j, y, ch);
                    }
                } else {
#pragma unroll
                    for (int ch = 0; ch < 3; ++ch) static_cast<P*>(J.res)[ch * J.rplane + (long)y * J.rpitch + j] = (P)acc[ch];
                }
            }
            continue;
        }
        if (j >= cw) continue;
        const int X = 2 * j;
        const int hdA = expand_hsum(X, cw, false, true), hdB = expand_hsum(X + 1, cw, false, true);
        const int iend = min(chh, ybase + rows);
        for (int i = ybase + ly; i < iend; i += 4) {
            const int Y = 2 * i;
            const int v0 = expand_vsum(Y, chh), v1 = expand_vsum(Y + 1, chh);
            const int dv[4] = {hdA * v0, hdB * v0, hdA * v1, hdB * v1};
            bool live[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) live[q] = X + (q & 1) < fw && Y + (q >> 1) < fh;
            S acc[4][3];
            for (int m = 0; m < J.K; ++m) {
                const BPlane cs = J.coarse[m];
                const BPlane fs = J.fine[m];
                const BPlane mk = J.mask[m];
                double coeff[4];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    coeff[q] = live[q] ? (double)static_cast<const M*>(mk.p)[(long)(Y + (q >> 1)) * mk.pitch + X + (q & 1)] / maxpyr : 0.0;
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) {
                    S o[4];
                    expand_block<P>([&](int y, int x) { return PlaneLoad<P>::get(cs, y, x, ch, 0); }, i, j, cw, chh, o);
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        if (!live[q]) continue;
                        // Laplacian value (src/pyramid.h:1705-1707), mask multiply
                        // (src/enfuse.h:601-620) and accumulation (:1576-1588).
                        const P lapv = Px::wrap(A::sub(FINE::get(fs, Y + (q >> 1), X + (q & 1), ch, frac), A::div(o[q], dv[q])));
                        const S v = Px::load(from_real<P>((double)lapv * coeff[q]));
                        acc[q][ch] = m == 0 ? v : Px::load(Px::wrap(A::add(v, acc[q][ch])));
                    }
                }
            }
            // Collapse step: + expand(result[l+1]), saturating (src/pyramid.h:1837-1886).
            const P* rc = static_cast<const P*>(J.rcoarse);
            P outv[4][3];
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                S o[4];
                expand_block<P>([&](int y, int x) { return Px::load(rc[ch * J.rcplane + (long)y * J.rcpitch + x]); }, i, j, cw, chh, o);
#pragma unroll
                for (int q = 0; q < 4; ++q) outv[q][ch] = live[q] ? Px::sat(A::add(acc[q][ch], A::div(o[q], dv[q]))) : (P)0;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (!live[q]) continue;
                const int xx = X + (q & 1), yy = Y + (q >> 1);
                if (LEVEL0) {
                    const uint8_t ua = J.ualpha[(long)yy * J.upitch + xx];
                    J.oalpha[(long)yy * J.oapitch + xx] = ua;
                    if (ua) {
                        I* o = static_cast<I*>(J.out) + (long)yy * J.opitch + 3L * xx;
#pragma unroll
                        for (int ch = 0; ch < 3; ++ch) o[ch] = to_image<I>(outv[q][ch], frac, dither, J.seed, xx, yy, ch);
                    }
                } else {
#pragma unroll
                    for (int ch = 0; ch < 3; ++ch) static_cast<P*>(J.res)[ch * J.rplane + (long)yy * J.rpitch + xx] = outv[q][ch];
                }
            }
        }
    }
}

}  // namespace

int launch_batch_weights(int itype, const BWeightJob* jobs, const int* tile0, int njobs, int total, const float* lut,
                         double w_exposure, double w_saturation, double scale, int rows, int grid, hipStream_t s)
{
    if (total <= 0) return 0;
    if (itype == ENBL_U8)
        hipLaunchKernelGGL((batch_weights_kernel<uint8_t, int16_t>), dim3(grid), dim3(256), 0, s, jobs, tile0, njobs, total, lut,
                           w_exposure, w_saturation, scale, rows);
    else if (itype == ENBL_U16)
        hipLaunchKernelGGL((batch_weights_kernel<uint16_t, int32_t>), dim3(grid), dim3(256), 0, s, jobs, tile0, njobs, total, lut,
                           w_exposure, w_saturation, scale, rows);
    else return ENBL_E_UNSUPPORTED;
    return 0;
}

int launch_batch_reduce(int itype, int frac, const BReduceJob* jobs, const int* tile0, int njobs, int total, int rows,
                        int grid, hipStream_t s)
{
    if (total <= 0) return 0;
    if (itype == ENBL_U8)
        hipLaunchKernelGGL((batch_reduce_kernel<int16_t, uint8_t>), dim3(grid), dim3(256), 0, s, jobs, tile0, njobs, total, rows, frac);
    else if (itype == ENBL_U16)
        hipLaunchKernelGGL((batch_reduce_kernel<int32_t, uint16_t>), dim3(grid), dim3(256), 0, s, jobs, tile0, njobs, total, rows, frac);
    else return ENBL_E_UNSUPPORTED;
    return 0;
}

int launch_batch_fuse(int itype, int frac, int level0, const BFuseJob* jobs, const int* tile0, int njobs, int total,
                      double maxpyr, int dither, int rows, int grid, hipStream_t s)
{
    if (total <= 0) return 0;
#define ENBL_BFUSE(P, M, I)                                                                                          \
    do {                                                                                                             \
        if (level0) hipLaunchKernelGGL((batch_fuse_kernel<P, M, I, true>), dim3(grid), dim3(256), 0, s, jobs, tile0, njobs, total, \
                                       maxpyr, frac, dither, rows);                                                  \
        else hipLaunchKernelGGL((batch_fuse_kernel<P, M, I, false>), dim3(grid), dim3(256), 0, s, jobs, tile0, njobs, total,   \
                                maxpyr, frac, dither, rows);                                                         \
    } while (0)
    if (itype == ENBL_U8) ENBL_BFUSE(int16_t, int16_t, uint8_t);
    else if (itype == ENBL_U16) ENBL_BFUSE(int32_t, int32_t, uint16_t);
    else return ENBL_E_UNSUPPORTED;
#undef ENBL_BFUSE
    return 0;
}

}  // namespace enbl
