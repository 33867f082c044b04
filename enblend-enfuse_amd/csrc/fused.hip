// Fused pyramid passes for gfx950.
//
// laplacian_level_kernel: one streaming pass over a Gaussian level G_l (or,
//   for level 0, over the interleaved input image, converted on the fly as
//   copyToPyramidImage does) that writes both the Laplacian level
//   L_l = G_l - expand(G_{l+1}) and the next Gaussian level G_{l+1} with its
//   alpha.  This is the reference's reduce (src/pyramid.h:166-573) followed by
//   its expand-subtract (src/pyramid.h:1102-1419, :1705-1707) with every
//   operation in the same order; the saving is that G_l is read once and the
//   intermediate G_{l+1} never makes an extra round trip through HBM.
//
// collapse_level_kernel: one streaming pass that blends the white and black
//   Laplacian levels (CartesianBlendFunctor, src/blend.h:41-76), adds the
//   expansion of the already collapsed coarser level (src/pyramid.h:1757-1784)
//   and, at level 0, converts straight into the output image with the
//   dithered copyFromPyramidImageIf (src/fixmath.h:2157-2294) under the union
//   alpha, which it also writes (src/enblend.h:461-463).
//
// Wave layout: lane L of a wave owns coarse column j = 60 * wave + L - 2 for
// L in [2, 61]; lanes 0, 1, 62, 63 compute the horizontal halo that the
// owned lanes read through cross-lane shuffles.  No wraparound here (the
// unfused kernels handle it).
#include <math.h>

#include <type_traits>

#include "enbl_device.h"
#include "enbl_fused.h"
#include "enbl_internal.h"
#include "pyr_common.h"

namespace enbl {
namespace {

constexpr int kOwnedF = kWave - 4;

// ---- sources of the fine level -------------------------------------------
template <class P> struct PlaneSrc {
    typedef typename Pix<P>::S S;
    template <int C>
    __device__ static __forceinline__ void load(const FusedLapArgs& a, int r, int c, bool both, S (&x0)[C], S (&x1)[C]) {
        const P* src = static_cast<const P*>(a.src);
#pragma unroll
        for (int ch = 0; ch < C; ++ch) Vec2<P>::load(src + ch * a.splane + (long)r * a.spitch + c, both, x0[ch], x1[ch]);
    }
};

// Interleaved integer image, v << frac (src/fixmath.h:197-202).
template <class I, class P> struct IntImageSrc {
    typedef typename Pix<P>::S S;
    template <int C>
    __device__ static __forceinline__ void load(const FusedLapArgs& a, int r, int c, bool both, S (&x0)[C], S (&x1)[C]) {
        const I* p = static_cast<const I*>(a.src) + (long)r * a.spitch + (long)c * C;
        I v[2 * C];
        if constexpr (sizeof(I) == 2 && C == 3) {
            // two RGB16 pixels = 12 bytes: one dwordx3 access when aligned
            if (both && ((uintptr_t)p & 3) == 0) {
                const uint3 d = *reinterpret_cast<const uint3*>(p);
                v[0] = (I)(d.x & 0xffffu); v[1] = (I)(d.x >> 16);
                v[2] = (I)(d.y & 0xffffu); v[3] = (I)(d.y >> 16);
                v[4] = (I)(d.z & 0xffffu); v[5] = (I)(d.z >> 16);
            } else {
#pragma unroll
                for (int k = 0; k < 2 * C; ++k) v[k] = (k < C || both) ? p[k] : (I)0;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 2 * C; ++k) v[k] = (k < C || both) ? p[k] : (I)0;
        }
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            x0[ch] = (S)(P)((P)v[ch] << a.frac);
            x1[ch] = both ? (S)(P)((P)v[C + ch] << a.frac) : (S)0;
        }
    }
};

// Interleaved float image, 1 + log1p(v) or 1 / (1 - v) (src/fixmath.h:165-186).
struct FloatImageSrc {
    typedef double S;
    __device__ static __forceinline__ double conv(float f) {
        const double v = (double)f;
        return v >= 0.0 ? 1.0 + log1p(v) : 1.0 / (1.0 - v);
    }
    template <int C>
    __device__ static __forceinline__ void load(const FusedLapArgs& a, int r, int c, bool both, S (&x0)[C], S (&x1)[C]) {
        const float* p = static_cast<const float*>(a.src) + (long)r * a.spitch + (long)c * C;
        float f0[C], f1[C];
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            f0[ch] = p[ch];
            f1[ch] = both ? p[C + ch] : 0.f;
        }
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            x0[ch] = conv(f0[ch]);
            x1[ch] = both ? conv(f1[ch]) : 0.0;
        }
    }
};

template <class S, int C>
struct FRow {
    S x0[C], x1[C];   // raw samples 2j, 2j+1 (unmasked)
    int m0, m1;       // alpha flags
};

template <class S, int C>
struct CoarseRow {
    S hA[C], hB[C];   // polyphase row sums of one collapsed/Gaussian coarse row
};

template <class P, int C, class SRC>
__global__ __launch_bounds__(256) void laplacian_level_kernel(FusedLapArgs a)
{
    typedef Pix<P> Px;
    typedef typename Px::A A;
    typedef typename Px::S S;
    typedef FRow<S, C> Row;
    typedef CoarseRow<S, C> CRow;

    const int lane = threadIdx.x & (kWave - 1);
    const int wx = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
    if (wx * kOwnedF - 2 >= a.cw) return;
    const int i0 = blockIdx.y * a.rows;
    if (i0 >= a.ch) return;
    const int i1 = min(a.ch, i0 + a.rows);
    const int fw = a.fw, fh = a.fh, cw = a.cw, chh = a.ch;
    const int j = wx * kOwnedF + lane - 2;
    const int c = 2 * j;
    const bool valid = j >= 0 && j < cw;
    const bool owned = lane >= 2 && lane < kWave - 2 && valid;
    const bool both = c + 1 < fw;
    const int rcase = c == fw - 1 ? 1 : (c == fw - 2 ? 2 : 0);

    auto load = [&](int r, Row& R) {
        R.m0 = 0;
        R.m1 = 0;
#pragma unroll
        for (int ch = 0; ch < C; ++ch) { R.x0[ch] = A::zero(); R.x1[ch] = A::zero(); }
        if (!valid) return;
        const uint8_t* ar = a.salpha + (long)r * a.sapitch;
        R.m0 = ar[c] ? 1 : 0;
        R.m1 = (both && ar[c + 1]) ? 1 : 0;
        SRC::template load<C>(a, r, c, both, R.x0, R.x1);
    };

    // Horizontal 5-tap sums with zero-extension (reduce with alpha).
    auto sums = [&](const Row& R, S (&h)[C], int& hm) {
        const int ml0 = lane_up(R.m0), ml1 = lane_up(R.m1), mr0 = lane_down(R.m0);
        const int qa = c >= 2 ? ml0 : 0, qb = c >= 1 ? ml1 : 0;
        hm = qa + 4 * qb + 6 * R.m0;
        if (rcase != 1) hm += 4 * R.m1;
        if (rcase == 0) hm += mr0;
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            const S x0 = R.m0 ? R.x0[ch] : A::zero();
            const S x1 = R.m1 ? R.x1[ch] : A::zero();
            const S l0 = lane_up(x0), l1 = lane_up(x1), r0 = lane_down(x0);
            const S va = c >= 2 ? l0 : A::zero();
            const S vb = c >= 1 ? l1 : A::zero();
            const S a0 = A::add(va, A::mul(vb, 4));
            if (rcase == 1) h[ch] = A::add(a0, A::mul(x0, 6));
            else if (rcase == 2) h[ch] = A::add(A::add(a0, A::mul(x0, 6)), A::mul(x1, 4));
            else h[ch] = A::add(A::add(A::add(a0, A::mul(x0, 6)), A::mul(x1, 4)), r0);
        }
    };

    // Polyphase sums of a coarse row held one value per lane (no wraparound).
    auto coarse_sums = [&](const S (&g)[C], CRow& out) {
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            const S x = g[ch];
            const S l = lane_up(x), r = lane_down(x);
            if (j == 0) {
                out.hA[ch] = A::add(A::add(A::zero(), A::mul(x, 6)), r);
                out.hB[ch] = A::mul(A::add(x, r), 4);
            } else if (j == cw - 1) {
                out.hA[ch] = A::add(l, A::mul(x, 6));
                out.hB[ch] = A::mul(x, 4);
            } else {
                out.hA[ch] = A::add(A::add(l, A::mul(x, 6)), r);
                out.hB[ch] = A::mul(A::add(x, r), 4);
            }
        }
    };

    const bool fw_even = (fw & 1) == 0;
    const int X = 2 * j;
    const bool xb = X + 1 < fw;
    const int hdA = expand_hsum(X, cw, false, fw_even);
    const int hdB = expand_hsum(X + 1, cw, false, fw_even);
    P* lap = static_cast<P*>(a.lap);
    P* coarse = static_cast<P*>(a.coarse);

    // Laplacian rows 2i and 2i+1 from the raw fine rows and the coarse rows
    // i-1 (p), i (q) and i+1 (n, when `more`).
    auto emit = [&](int i, const CRow& p, const CRow& q, const CRow& n, bool more, const Row& RA, const Row& RB) {
        if (!owned) return;
        const int Y = 2 * i;
        const bool yb = Y + 1 < fh;
        const int v0 = expand_vsum(Y, chh), v1 = expand_vsum(Y + 1, chh);
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            S o00, o10, o01, o11;
            if (more) {
                o00 = A::add(A::add(p.hA[ch], A::mul(q.hA[ch], 6)), n.hA[ch]);
                o10 = A::add(A::add(p.hB[ch], A::mul(q.hB[ch], 6)), n.hB[ch]);
                o01 = A::add(q.hA[ch], n.hA[ch]);
                o11 = A::add(q.hB[ch], n.hB[ch]);
            } else {
                o00 = A::add(p.hA[ch], A::mul(q.hA[ch], 6));
                o10 = A::add(p.hB[ch], A::mul(q.hB[ch], 6));
                o01 = q.hA[ch];
                o11 = q.hB[ch];
            }
            P* row0 = lap + ch * a.lplane + (long)Y * a.lpitch + X;
            Vec2<P>::store(row0, xb, Px::wrap(A::sub(RA.x0[ch], A::div(o00, hdA * v0))),
                           Px::wrap(A::sub(RA.x1[ch], A::div(o10, hdB * v0))));
            if (yb)
                Vec2<P>::store(row0 + a.lpitch, xb, Px::wrap(A::sub(RB.x0[ch], A::div(o01, hdA * v1))),
                               Px::wrap(A::sub(RB.x1[ch], A::div(o11, hdB * v1))));
        }
    };

    const int k0 = max(i0 - 1, 0);
    const int k1 = min(i1, chh - 1);
    Row w[5], n0, n1;
    S h[5][C];
    int q[5];
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        q[t] = 0;
#pragma unroll
        for (int ch = 0; ch < C; ++ch) h[t][ch] = A::zero();
        const int r = 2 * k0 - 2 + t;
        if (r >= 0 && r < fh) load(r, w[t]);
    }
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const int r = 2 * k0 - 2 + t;
        if (r >= 0 && r < fh) sums(w[t], h[t], q[t]);
    }
    CRow g0, g1, g2;   // coarse rows k-2, k-1, k
#pragma unroll
    for (int ch = 0; ch < C; ++ch) {
        g0.hA[ch] = g0.hB[ch] = g1.hA[ch] = g1.hB[ch] = g2.hA[ch] = g2.hB[ch] = A::zero();
    }

    for (int k = k0; k <= k1; ++k) {
        const int r = 2 * k;
        if (k < k1) {
            if (r + 3 < fh) load(r + 3, n0);
            if (r + 4 < fh) load(r + 4, n1);
        }
        if (r + 1 < fh) sums(w[3], h[3], q[3]);
        if (r + 2 < fh) sums(w[4], h[4], q[4]);

        // Vertical reduce: coarse row k (src/pyramid.h:350-384, :459-504).
        const int q0 = r == 0 ? 0 : q[0] + 4 * q[1];
        int ap;
        if (r == fh - 1) ap = q0 + 6 * q[2];
        else if (r == fh - 2) ap = q0 + 6 * q[2] + 4 * q[3];
        else ap = q0 + 6 * q[2] + 4 * q[3] + q[4];
        S gk[C];
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            const S a0 = r == 0 ? A::zero() : A::add(h[0][ch], A::mul(h[1][ch], 4));
            S v;
            if (r == fh - 1) v = A::add(a0, A::mul(h[2][ch], 6));
            else if (r == fh - 2) v = A::add(A::add(a0, A::mul(h[2][ch], 6)), A::mul(h[3][ch], 4));
            else v = A::add(A::add(A::add(a0, A::mul(h[2][ch], 6)), A::mul(h[3][ch], 4)), h[4][ch]);
            const P out = ap ? Px::wrap(A::div(v, ap)) : (P)0;
            gk[ch] = Px::load(out);
            if (owned && k >= i0 && k < i1) coarse[ch * a.cplane + (long)k * a.cpitch + j] = out;
        }
        if (owned && k >= i0 && k < i1) a.calpha[(long)k * a.capitch + j] = ap ? 255 : 0;

        g0 = g1;
        g1 = g2;
        coarse_sums(gk, g2);

        if (k >= 1 && k - 1 >= i0 && k - 1 < i1) {
            CRow p = g0;
            if (k - 1 == 0) {
#pragma unroll
                for (int ch = 0; ch < C; ++ch) { p.hA[ch] = A::zero(); p.hB[ch] = A::zero(); }
            }
            emit(k - 1, p, g1, g2, true, w[0], w[1]);
        }
        if (k == chh - 1 && k >= i0 && k < i1) {
            CRow p = g1;
            if (k == 0) {
#pragma unroll
                for (int ch = 0; ch < C; ++ch) { p.hA[ch] = A::zero(); p.hB[ch] = A::zero(); }
            }
            emit(k, p, g2, g2, false, w[2], w[3]);
        }

        if (k < k1) {
            w[0] = w[2]; w[1] = w[3]; w[2] = w[4]; w[3] = n0; w[4] = n1;
#pragma unroll
            for (int ch = 0; ch < C; ++ch) { h[0][ch] = h[2][ch]; h[1][ch] = h[3][ch]; h[2][ch] = h[4][ch]; }
            q[0] = q[2]; q[1] = q[3]; q[2] = q[4];
        }
    }
}

// ---- fused blend + collapse (+ output) -----------------------------------
template <class S, int C, class M>
struct FinePair {
    S b0[C], b1[C], w0[C], w1[C];   // black / white samples X, X+1 of one fine row
    M m0, m1;                       // mask samples
    uint8_t wa0, wa1, ba0, ba1;     // alphas (output level only)
};

template <class P> __device__ __forceinline__ P blend_px(double wc, P w, P b) {
    // CartesianBlendFunctor (src/blend.h:47-70)
    if (wc >= 1.0) return w;
    if (wc <= 0.0) return b;
    return from_real<P>((wc * (double)w) + ((1.0 - wc) * (double)b));
}

template <class P, class I> struct OutConv;
template <class P> struct OutConv<P, uint8_t> {
    __device__ static __forceinline__ uint8_t conv(P v, int frac, int dither, uint64_t seed, int x, int y, int c) {
        const int32_t q = (int32_t)v >> frac, f = (int32_t)v & ((1 << frac) - 1);
        int32_t r;
        if (f >= (1 << (frac - 2)) && f < (3 << (frac - 2)))
            r = dither == ENBL_DITHER_DET ? (dither_rand(seed, x, y, c, frac) <= f ? q + 1 : q) : (f >= (1 << (frac - 1)) ? q + 1 : q);
        else r = f >= (1 << (frac - 2)) ? q + 1 : q;
        return (uint8_t)min(255, max(0, r));
    }
};
template <class P> struct OutConv<P, uint16_t> {
    __device__ static __forceinline__ uint16_t conv(P v, int frac, int dither, uint64_t seed, int x, int y, int c) {
        const int32_t q = (int32_t)v >> frac, f = (int32_t)v & ((1 << frac) - 1);
        int32_t r;
        if (f >= (1 << (frac - 2)) && f < (3 << (frac - 2)))
            r = dither == ENBL_DITHER_DET ? (dither_rand(seed, x, y, c, frac) <= f ? q + 1 : q) : (f >= (1 << (frac - 1)) ? q + 1 : q);
        else r = f >= (1 << (frac - 2)) ? q + 1 : q;
        return (uint16_t)min(65535, max(0, r));
    }
};
template <> struct OutConv<double, float> {
    __device__ static __forceinline__ float conv(double v, int, int, uint64_t, int, int, int) {
        return (float)(v >= 1.0 ? expm1(v - 1.0) : 1.0 - 1.0 / v);
    }
};

// I = void: write the collapsed level back into black (levels >= 1);
// otherwise convert into the interleaved output image of type I.
template <class P, class M, int C, class I, bool MASK8>
__global__ __launch_bounds__(256) void collapse_level_kernel(FusedCollapseArgs a)
{
    typedef Pix<P> Px;
    typedef typename Px::A A;
    typedef typename Px::S S;
    typedef FinePair<S, C, M> Fine;
    constexpr bool OUT = !std::is_same<I, void>::value;

    const int lane = threadIdx.x & (kWave - 1);
    const int wx = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
    if (wx * kOwned - 1 >= a.cw) return;
    const int i0 = blockIdx.y * a.rows;
    if (i0 >= a.ch) return;
    const int i1 = min(a.ch, i0 + a.rows);
    const int cw = a.cw, chh = a.ch, fw = a.fw, fh = a.fh;
    const int j = wx * kOwned + lane - 1;
    const bool valid = j >= 0 && j < cw;
    const bool owned = lane >= 1 && lane <= kOwned && valid;
    const int X = 2 * j;
    const bool xb = X + 1 < fw;
    const bool fw_even = (fw & 1) == 0;
    const int hdA = expand_hsum(X, cw, false, fw_even);
    const int hdB = expand_hsum(X + 1, cw, false, fw_even);
    const P* cs = static_cast<const P*>(a.coarse);
    P* blk = static_cast<P*>(a.black);
    const P* wht = static_cast<const P*>(a.white);

    auto load_coarse = [&](int i, S (&x)[C]) {
#pragma unroll
        for (int ch = 0; ch < C; ++ch) x[ch] = valid ? Px::load(cs[ch * a.cplane + (long)i * a.cpitch + j]) : A::zero();
    };
    auto coarse_sums = [&](const S (&x)[C], S (&hA)[C], S (&hB)[C]) {
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            const S v = x[ch];
            const S l = lane_up(v), r = lane_down(v);
            if (j == 0) {
                hA[ch] = A::add(A::add(A::zero(), A::mul(v, 6)), r);
                hB[ch] = A::mul(A::add(v, r), 4);
            } else if (j == cw - 1) {
                hA[ch] = A::add(l, A::mul(v, 6));
                hB[ch] = A::mul(v, 4);
            } else {
                hA[ch] = A::add(A::add(l, A::mul(v, 6)), r);
                hB[ch] = A::mul(A::add(v, r), 4);
            }
        }
    };
    auto load_fine = [&](int Y, Fine& F) {
        if (!owned || Y >= fh) return;
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            Vec2<P>::load(blk + ch * a.fplane + (long)Y * a.fpitch + X, xb, F.b0[ch], F.b1[ch]);
            Vec2<P>::load(wht + ch * a.wplane + (long)Y * a.wpitch + X, xb, F.w0[ch], F.w1[ch]);
        }
        if (MASK8) {
            const uint8_t* m = static_cast<const uint8_t*>(a.mask) + (long)Y * a.mpitch + X;
            F.m0 = (M)((M)m[0] << a.mfrac);
            F.m1 = xb ? (M)((M)m[1] << a.mfrac) : (M)0;
        } else {
            const M* m = static_cast<const M*>(a.mask) + (long)Y * a.mpitch + X;
            F.m0 = m[0];
            F.m1 = xb ? m[1] : (M)0;
        }
        if (OUT) {
            F.wa0 = a.walpha[(long)Y * a.wapitch + X];
            F.ba0 = a.balpha[(long)Y * a.bapitch + X];
            F.wa1 = xb ? a.walpha[(long)Y * a.wapitch + X + 1] : 0;
            F.ba1 = xb ? a.balpha[(long)Y * a.bapitch + X + 1] : 0;
        }
    };
    auto finish = [&](int Y, const Fine& F, const S (&e0)[C], const S (&e1)[C]) {
        const double wc0 = (double)F.m0 / a.white_value;
        const double wc1 = (double)F.m1 / a.white_value;
        P o0[C], o1[C];
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            o0[ch] = Px::sat(A::add(Px::load(blend_px<P>(wc0, (P)F.w0[ch], (P)F.b0[ch])), e0[ch]));
            o1[ch] = Px::sat(A::add(Px::load(blend_px<P>(wc1, (P)F.w1[ch], (P)F.b1[ch])), e1[ch]));
        }
        if constexpr (!OUT) {
#pragma unroll
            for (int ch = 0; ch < C; ++ch) Vec2<P>::store(blk + ch * a.fplane + (long)Y * a.fpitch + X, xb, o0[ch], o1[ch]);
        } else {
            typedef typename std::conditional<OUT, I, uint8_t>::type IT;
            IT* img = static_cast<IT*>(a.img) + (long)Y * a.ipitch + (long)X * C;
            const bool u0 = F.wa0 || F.ba0;
            const bool u1 = F.wa1 || F.ba1;
            if constexpr (sizeof(IT) == 2 && C == 3) {
                if (u0 && xb && u1 && ((uintptr_t)img & 3) == 0) {
                    IT v[6];
#pragma unroll
                    for (int ch = 0; ch < 3; ++ch) {
                        v[ch] = OutConv<P, IT>::conv(o0[ch], a.ifrac, a.dither, a.seed, X, Y, ch);
                        v[3 + ch] = OutConv<P, IT>::conv(o1[ch], a.ifrac, a.dither, a.seed, X + 1, Y, ch);
                    }
                    *reinterpret_cast<uint3*>(img) = make_uint3((uint32_t)v[0] | ((uint32_t)v[1] << 16),
                                                                (uint32_t)v[2] | ((uint32_t)v[3] << 16),
                                                                (uint32_t)v[4] | ((uint32_t)v[5] << 16));
                    if (F.wa0) a.balpha[(long)Y * a.bapitch + X] = 255;
                    if (F.wa1) a.balpha[(long)Y * a.bapitch + X + 1] = 255;
                    return;
                }
            }
            if (u0) {
#pragma unroll
                for (int ch = 0; ch < C; ++ch) img[ch] = OutConv<P, IT>::conv(o0[ch], a.ifrac, a.dither, a.seed, X, Y, ch);
            }
            if (xb && u1) {
#pragma unroll
                for (int ch = 0; ch < C; ++ch) img[C + ch] = OutConv<P, IT>::conv(o1[ch], a.ifrac, a.dither, a.seed, X + 1, Y, ch);
            }
            if (F.wa0) a.balpha[(long)Y * a.bapitch + X] = 255;
            if (xb && F.wa1) a.balpha[(long)Y * a.bapitch + X + 1] = 255;
        }
    };

    S pA[C], pB[C], cA[C], cB[C], nA[C], nB[C], xr[C], xn[C];
#pragma unroll
    for (int ch = 0; ch < C; ++ch) { pA[ch] = pB[ch] = cA[ch] = cB[ch] = nA[ch] = nB[ch] = xr[ch] = xn[ch] = A::zero(); }
    {
        S t[C];
        if (i0 > 0) {
            load_coarse(i0 - 1, t);
            coarse_sums(t, pA, pB);
        }
        load_coarse(i0, t);
        coarse_sums(t, cA, cB);
    }
    if (i0 + 1 < chh) load_coarse(i0 + 1, xr);
    Fine f0, f1, g0, g1;
    load_fine(2 * i0, f0);
    load_fine(2 * i0 + 1, f1);

    for (int i = i0; i < i1; ++i) {
        const bool more = i + 1 < chh;
        if (i + 1 < i1) {
            if (i + 2 < chh) load_coarse(i + 2, xn);
            load_fine(2 * i + 2, g0);
            load_fine(2 * i + 3, g1);
        }
        if (more) coarse_sums(xr, nA, nB);
        const int Y = 2 * i;
        const int v0 = expand_vsum(Y, chh), v1 = expand_vsum(Y + 1, chh);
        S e00[C], e10[C], e01[C], e11[C];
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            S o00, o10, o01, o11;
            if (more) {
                o00 = A::add(A::add(pA[ch], A::mul(cA[ch], 6)), nA[ch]);
                o10 = A::add(A::add(pB[ch], A::mul(cB[ch], 6)), nB[ch]);
                o01 = A::add(cA[ch], nA[ch]);
                o11 = A::add(cB[ch], nB[ch]);
            } else {
                o00 = A::add(pA[ch], A::mul(cA[ch], 6));
                o10 = A::add(pB[ch], A::mul(cB[ch], 6));
                o01 = cA[ch];
                o11 = cB[ch];
            }
            e00[ch] = A::div(o00, hdA * v0);
            e10[ch] = A::div(o10, hdB * v0);
            e01[ch] = A::div(o01, hdA * v1);
            e11[ch] = A::div(o11, hdB * v1);
        }
        if (owned) {
            finish(Y, f0, e00, e10);
            if (Y + 1 < fh) finish(Y + 1, f1, e01, e11);
        }
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            pA[ch] = cA[ch]; pB[ch] = cB[ch]; cA[ch] = nA[ch]; cB[ch] = nB[ch]; xr[ch] = xn[ch];
        }
        f0 = g0;
        f1 = g1;
    }
}

inline int strip_rows_f(int waves_x, int out_rows)
{
    const long target = 8192;
    long r = ((long)waves_x * out_rows + target - 1) / target;
    if (r < 4) r = 4;
    if (r > 32) r = 32;
    return (int)r;
}

template <class P, class SRC>
void launch_lap_t(FusedLapArgs a, int C, hipStream_t s)
{
    const int waves_x = (a.cw + kOwnedF - 1) / kOwnedF + 1;
    if (a.rows <= 0) a.rows = strip_rows_f(waves_x, a.ch);
    dim3 grid((waves_x + 3) / 4, (a.ch + a.rows - 1) / a.rows);
    if (C == 1) hipLaunchKernelGGL((laplacian_level_kernel<P, 1, SRC>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((laplacian_level_kernel<P, 3, SRC>), grid, dim3(256), 0, s, a);
}

template <class P, class M, class I, bool MASK8>
void launch_col_t(FusedCollapseArgs a, int C, hipStream_t s)
{
    const int waves_x = (a.cw + kOwned - 1) / kOwned + 1;
    if (a.rows <= 0) a.rows = strip_rows_f(waves_x, a.ch);
    dim3 grid((waves_x + 3) / 4, (a.ch + a.rows - 1) / a.rows);
    if (C == 1) hipLaunchKernelGGL((collapse_level_kernel<P, M, 1, I, MASK8>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((collapse_level_kernel<P, M, 3, I, MASK8>), grid, dim3(256), 0, s, a);
}

}  // namespace

int launch_laplacian_level(int ptype, int src_kind, const FusedLapArgs& a, int C, hipStream_t s)
{
    if (C != 1 && C != 3) return ENBL_E_UNSUPPORTED;
    if (src_kind == 0) {
        if (ptype == ENBL_I16) { launch_lap_t<int16_t, PlaneSrc<int16_t>>(a, C, s); return 0; }
        if (ptype == ENBL_I32) { launch_lap_t<int32_t, PlaneSrc<int32_t>>(a, C, s); return 0; }
        if (ptype == ENBL_F64) { launch_lap_t<double, PlaneSrc<double>>(a, C, s); return 0; }
    } else if (src_kind == ENBL_U8 && ptype == ENBL_I16) {
        launch_lap_t<int16_t, IntImageSrc<uint8_t, int16_t>>(a, C, s); return 0;
    } else if (src_kind == ENBL_U16 && ptype == ENBL_I32) {
        launch_lap_t<int32_t, IntImageSrc<uint16_t, int32_t>>(a, C, s); return 0;
    } else if (src_kind == ENBL_F32 && ptype == ENBL_F64) {
        launch_lap_t<double, FloatImageSrc>(a, C, s); return 0;
    }
    return ENBL_E_UNSUPPORTED;
}

int launch_collapse_level(int ptype, int mtype, int out_itype, bool mask8, const FusedCollapseArgs& a, int C, hipStream_t s)
{
    if (C != 1 && C != 3) return ENBL_E_UNSUPPORTED;
#define ENBL_COL(P, M)                                                                                      \
    do {                                                                                                    \
        if (out_itype == 0) { if (mask8) launch_col_t<P, M, void, true>(a, C, s); else launch_col_t<P, M, void, false>(a, C, s); return 0; } \
        if (out_itype == ENBL_U8 && std::is_integral<P>::value) { launch_col_t<P, M, uint8_t, true>(a, C, s); return 0; }       \
        if (out_itype == ENBL_U16 && std::is_integral<P>::value) { launch_col_t<P, M, uint16_t, true>(a, C, s); return 0; }     \
    } while (0)
    if (ptype == ENBL_I16 && mtype == ENBL_I16) ENBL_COL(int16_t, int16_t);
    if (ptype == ENBL_I32 && mtype == ENBL_I32) ENBL_COL(int32_t, int32_t);
    if (ptype == ENBL_F64 && mtype == ENBL_I32) {
        if (out_itype == 0) { if (mask8) launch_col_t<double, int32_t, void, true>(a, C, s); else launch_col_t<double, int32_t, void, false>(a, C, s); return 0; }
        if (out_itype == ENBL_F32) { launch_col_t<double, int32_t, float, true>(a, C, s); return 0; }
    }
#undef ENBL_COL
    return ENBL_E_UNSUPPORTED;
}

}  // namespace enbl
