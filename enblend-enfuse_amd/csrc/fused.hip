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

#include <algorithm>
#include <type_traits>

#include "enbl_device.h"
#include "enbl_fused.h"
#include "enbl_internal.h"
#include "pyr_common.h"
#include "row_fetch.h"

namespace enbl {
namespace {

constexpr int kOwnedF = kWave - 4;

template <class S, int C>
struct FRow {
    S x0[C], x1[C];   // raw samples 2j, 2j+1 (unmasked)
    int m0, m1;       // alpha flags
};

template <class S, int C>
struct CoarseRow {
    S hA[C], hB[C];   // polyphase row sums of one collapsed/Gaussian coarse row
};

// Streaming structure: coarse row k of the strip needs fine rows 2k-2..2k+2
// (horizontal sums of rows 2k+1, 2k+2 are formed at step k) and the
// Laplacian of fine rows 2k-2, 2k-1 is written at step k, once coarse row k
// exists.  Raw fine rows live in a ring of 8 register slots (row r in slot
// (r - 2k0 + 2) & 7); the loop is unrolled by its period of 4 steps so every
// slot index is a constant and no in-flight register is ever copied.  Step k
// loads rows 2k+3, 2k+4 for step k+1 and issues the same memory instructions
// whatever it writes (rows and lanes that must not be touched get a zero-size
// resource or an out-of-range offset), so the wait for the rows of step k-1
// counts exactly and never waits for a store.
template <class SRC> struct ConvertsOnUnpack : std::false_type {};
template <int C> struct ConvertsOnUnpack<FloatImageSrc<C>> : std::true_type {};

template <class P, int C, class SRC, bool AW>
__global__ __launch_bounds__(256) void laplacian_level_kernel(FusedLapArgs a)
{
    typedef Pix<P> Px;
    typedef typename Px::A A;
    typedef typename Px::S S;
    typedef FRow<S, C> Row;
    typedef CoarseRow<S, C> CRow;
    typedef BPair<P> B;
    constexpr bool INT = std::is_same<A, AccI>::value;
    constexpr uint32_t ES = sizeof(P);
    struct RawRow {
        typename SRC::Raw px;
        AlphaPair<AW> al;
    };

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
    LaneCols L;
    L.valid = valid;
    L.both = both;
    L.tailw = !wave_all(!valid || both);
    L.c = c;

    auto fetch = [&](int r, bool want, RawRow& R) {
        const bool ok = want && r >= 0 && r < fh;
        const int rr = ok ? r : 0;
        SRC::fetch(a, fw, rr, ok, L, R.px);
        R.al.fetch(a.salpha + (long)rr * a.sapitch, ok, fw, L);
    };
    auto unpack = [&](const RawRow& R, Row& X) {
        SRC::unpack(a, R.px, X.x0, X.x1);
        X.m0 = valid && R.al.first() ? 1 : 0;
        X.m1 = valid && both && R.al.second() ? 1 : 0;
    };

    // Horizontal 5-tap sums with zero-extension (reduce with alpha); the
    // right-border cases are selected, not branched, per lane.
    auto sums = [&](const Row& R, S (&h)[C], int& hm) {
        const int ml0 = lane_up(R.m0), ml1 = lane_up(R.m1), mr0 = lane_down(R.m0);
        const int qa = c >= 2 ? ml0 : 0, qb = c >= 1 ? ml1 : 0;
        hm = qa + 4 * qb + 6 * R.m0 + (rcase != 1 ? 4 * R.m1 : 0) + (rcase == 0 ? mr0 : 0);
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            const S x0 = R.m0 ? R.x0[ch] : A::zero();
            const S x1 = R.m1 ? R.x1[ch] : A::zero();
            const S l0 = lane_up(x0), l1 = lane_up(x1), r0 = lane_down(x0);
            const S va = c >= 2 ? l0 : A::zero();
            const S vb = c >= 1 ? l1 : A::zero();
            const S a0 = A::add(va, A::mul(vb, 4));
            const S t1 = A::add(a0, A::mul(x0, 6));
            const S t2 = A::add(t1, A::mul(x1, 4));
            const S t0 = A::add(t2, r0);
            h[ch] = rcase == 1 ? t1 : (rcase == 2 ? t2 : t0);
        }
    };

    // Polyphase sums of a coarse row held one value per lane (no wraparound).
    const bool cfirst = j == 0, clast = j == cw - 1 && j != 0;
    auto coarse_sums = [&](const S (&g)[C], CRow& out) {
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            const S x = g[ch];
            const S l = lane_up(x), r = lane_down(x);
            const S t = A::add(cfirst ? A::zero() : l, A::mul(x, 6));
            out.hA[ch] = clast ? t : A::add(t, r);
            out.hB[ch] = clast ? A::mul(x, 4) : A::mul(A::add(x, r), 4);
        }
    };

    const bool fw_even = (fw & 1) == 0;
    const int X = 2 * j;
    const bool xb = X + 1 < fw;
    const int hdA = expand_hsum(X, cw, false, fw_even);
    const int hdB = expand_hsum(X + 1, cw, false, fw_even);
    // Power-of-two weight sums (every column but the first and last) divide
    // by a shift; the decision is per wave and row, so it is a uniform branch.
    const bool colp2 = wave_all(!owned || (((hdA & (hdA - 1)) == 0) && ((hdB & (hdB - 1)) == 0)));
    const int sA = __builtin_ctz((unsigned)hdA), sB = __builtin_ctz((unsigned)hdB);
    const uint32_t off2 = owned && xb ? (uint32_t)X * ES : kBufSkip;    // pair stores
    const uint32_t off1 = owned && !xb ? (uint32_t)X * ES : kBufSkip;   // last odd column
    const uint32_t offc = owned ? (uint32_t)j * ES : kBufSkip;          // coarse level
    const uint32_t offa = owned ? (uint32_t)j : kBufSkip;               // coarse alpha
    const uint32_t lbytes = (uint32_t)fw * ES;
    P* lap = static_cast<P*>(a.lap);
    P* coarse = static_cast<P*>(a.coarse);

    // Laplacian rows 2i and 2i+1 from the raw fine rows and the coarse rows
    // i-1 (p), i (q) and i+1 (n, when `more`); rows outside the strip are
    // computed and dropped (`ok` false).
    auto emit = [&](int i, const CRow& p, const CRow& q, const CRow& n, bool more, const RawRow& RA, const RawRow& RB,
                    bool ok) {
        const int Y = 2 * i;
        const bool yb = Y + 1 < fh;
        const int v0 = expand_vsum(Y, chh), v1 = expand_vsum(Y + 1, chh);
        S xa0[C], xa1[C], xb0[C], xb1[C];
        SRC::unpack(a, RA.px, xa0, xa1);
        SRC::unpack(a, RB.px, xb0, xb1);
        const bool fast = INT && colp2 && (v0 & (v0 - 1)) == 0 && (v1 & (v1 - 1)) == 0;
        const int s0 = __builtin_ctz((unsigned)v0), s1 = __builtin_ctz((unsigned)v1);
        const P* lrow = lap + (long)Y * a.lpitch;
        const uint32_t b0 = ok ? lbytes : 0u, b1 = ok && yb ? lbytes : 0u;
        // one channel at a time: its four expansions, differences and stores
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
            S e00, e10, e01, e11;
            if constexpr (INT) {
                if (fast) {
                    e00 = div_pow2(o00, sA + s0);
                    e10 = div_pow2(o10, sB + s0);
                    e01 = div_pow2(o01, sA + s1);
                    e11 = div_pow2(o11, sB + s1);
                } else {
                    e00 = o00 / (hdA * v0);
                    e10 = o10 / (hdB * v0);
                    e01 = o01 / (hdA * v1);
                    e11 = o11 / (hdB * v1);
                }
            } else {
                e00 = A::div(o00, hdA * v0);
                e10 = A::div(o10, hdB * v0);
                e01 = A::div(o01, hdA * v1);
                e11 = A::div(o11, hdB * v1);
            }
            const __amdgpu_buffer_rsrc_t r0 = buf_rsrc(lrow + ch * a.lplane, b0);
            const __amdgpu_buffer_rsrc_t r1 = buf_rsrc(lrow + ch * a.lplane + a.lpitch, b1);
            const P y00 = Px::wrap(A::sub(xa0[ch], e00)), y10 = Px::wrap(A::sub(xa1[ch], e10));
            const P y01 = Px::wrap(A::sub(xb0[ch], e01)), y11 = Px::wrap(A::sub(xb1[ch], e11));
            B::store2(r0, off2, y00, y10);
            B::store2(r1, off2, y01, y11);
            if (L.tailw) {
                B::store1(r0, off1, y00);
                B::store1(r1, off1, y01);
            }
        }
    };

    const int k0 = max(i0 - 1, 0);
    const int k1 = min(i1, chh - 1);
    const int base = 2 * k0 - 2;
    RawRow ring[8];
    // Horizontal sums of the window rows 2k-2..2k+2: the two oldest only
    // enter as h(2k-2) + 4 h(2k-1), kept combined (hc, qc).
    S hc[C], h2[C], h3[C], h4[C];
    int qc = 0, q2 = 0, q3 = 0, q4 = 0;
#pragma unroll
    for (int t = 0; t < 5; ++t) fetch(base + t, true, ring[t]);
    {
        S h0[C], h1[C];
        int q0 = 0, q1 = 0;
#pragma unroll
        for (int ch = 0; ch < C; ++ch) h0[ch] = h1[ch] = h2[ch] = h3[ch] = h4[ch] = A::zero();
        Row R;
        if (base >= 0 && base < fh) { unpack(ring[0], R); sums(R, h0, q0); }
        if (base + 1 >= 0 && base + 1 < fh) { unpack(ring[1], R); sums(R, h1, q1); }
        if (base + 2 >= 0 && base + 2 < fh) { unpack(ring[2], R); sums(R, h2, q2); }
#pragma unroll
        for (int ch = 0; ch < C; ++ch) hc[ch] = A::add(h0[ch], A::mul(h1[ch], 4));
        qc = q0 + 4 * q1;
    }
    CRow g0, g1, g2;   // coarse rows k-2, k-1, k
#pragma unroll
    for (int ch = 0; ch < C; ++ch) {
        g0.hA[ch] = g0.hB[ch] = g1.hA[ch] = g1.hB[ch] = g2.hA[ch] = g2.hB[ch] = A::zero();
    }

    auto step = [&](auto ph, int k) {
        constexpr int PH = decltype(ph)::value;
        constexpr int S0 = (2 * PH) & 7, S1 = (2 * PH + 1) & 7, S2 = (2 * PH + 2) & 7, S3 = (2 * PH + 3) & 7;
        constexpr int S4 = (2 * PH + 4) & 7, S5 = (2 * PH + 5) & 7, S6 = (2 * PH + 6) & 7;
        const int r = 2 * k;
        fetch(r + 3, k < k1, ring[S5]);
        fetch(r + 4, k < k1, ring[S6]);
        if (r + 1 < fh) {
            Row R;
            unpack(ring[S3], R);
            sums(R, h3, q3);
        }
        if (r + 2 < fh) {
            Row R;
            unpack(ring[S4], R);
            sums(R, h4, q4);
        }

        // Vertical reduce: coarse row k (src/pyramid.h:350-384, :459-504).
        const int q0 = r == 0 ? 0 : qc;
        int ap;
        if (r == fh - 1) ap = q0 + 6 * q2;
        else if (r == fh - 2) ap = q0 + 6 * q2 + 4 * q3;
        else ap = q0 + 6 * q2 + 4 * q3 + q4;
        S v[C];
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            const S a0 = r == 0 ? A::zero() : hc[ch];
            if (r == fh - 1) v[ch] = A::add(a0, A::mul(h2[ch], 6));
            else if (r == fh - 2) v[ch] = A::add(A::add(a0, A::mul(h2[ch], 6)), A::mul(h3[ch], 4));
            else v[ch] = A::add(A::add(A::add(a0, A::mul(h2[ch], 6)), A::mul(h3[ch], 4)), h4[ch]);
        }
        P out[C];
        if constexpr (INT) {
            const int d = ap ? ap : 1;
            if (wave_all((d & (d - 1)) == 0)) {
                const int sd = __builtin_ctz((unsigned)d);
#pragma unroll
                for (int ch = 0; ch < C; ++ch) out[ch] = ap ? Px::wrap(div_pow2(v[ch], sd)) : (P)0;
            } else {
#pragma unroll
                for (int ch = 0; ch < C; ++ch) out[ch] = ap ? Px::wrap(v[ch] / d) : (P)0;
            }
        } else {
#pragma unroll
            for (int ch = 0; ch < C; ++ch) out[ch] = ap ? Px::wrap(A::div(v[ch], ap)) : (P)0;
        }
        S gk[C];
#pragma unroll
        for (int ch = 0; ch < C; ++ch) gk[ch] = Px::load(out[ch]);
        {
            const bool mine = k >= i0 && k < i1;
            const uint32_t cb = mine ? (uint32_t)cw * ES : 0u;
            const P* crow = coarse + (long)k * a.cpitch;
#pragma unroll
            for (int ch = 0; ch < C; ++ch) {
                const __amdgpu_buffer_rsrc_t rs = buf_rsrc(crow + ch * a.cplane, cb);
                if constexpr (sizeof(P) == 2) __builtin_amdgcn_raw_buffer_store_b16((unsigned short)out[ch], rs, offc, 0, 0);
                else if constexpr (sizeof(P) == 4) __builtin_amdgcn_raw_buffer_store_b32((uint32_t)out[ch], rs, offc, 0, 0);
                else B::store1(rs, offc, out[ch]);
            }
            __builtin_amdgcn_raw_buffer_store_b8((unsigned char)(ap ? 255 : 0),
                                                 buf_rsrc(a.calpha + (long)k * a.capitch, mine ? (uint32_t)cw : 0u), offa, 0, 0);
        }

        g0 = g1;
        g1 = g2;
        coarse_sums(gk, g2);

        {
            CRow p = g0;
            if (k - 1 == 0) {
#pragma unroll
                for (int ch = 0; ch < C; ++ch) { p.hA[ch] = A::zero(); p.hB[ch] = A::zero(); }
            }
            emit(k - 1, p, g1, g2, true, ring[S0], ring[S1], k >= 1 && k - 1 >= i0 && k - 1 < i1);
        }
        if (k == chh - 1 && k >= i0 && k < i1) {
            CRow p = g1;
            if (k == 0) {
#pragma unroll
                for (int ch = 0; ch < C; ++ch) { p.hA[ch] = A::zero(); p.hB[ch] = A::zero(); }
            }
            emit(k, p, g2, g2, false, ring[S2], ring[S3], true);
        }

#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            hc[ch] = A::add(h2[ch], A::mul(h3[ch], 4));
            h2[ch] = h4[ch];
        }
        qc = q2 + 4 * q3;
        q2 = q4;
    };

    if constexpr (ConvertsOnUnpack<SRC>::value) {
        // Float sources convert (log1p) on every unpack, so this pass is
        // arithmetic-bound: one step body, with the ring shifted by copies.
        for (int k = k0; k <= k1; ++k) {
            step(Phase<0>(), k);
#pragma unroll
            for (int t = 0; t < 5; ++t) ring[t] = ring[t + 2];
        }
    } else {
        int k = k0;
        for (; k + 3 <= k1; k += 4) {
            step(Phase<0>(), k);
            step(Phase<1>(), k + 1);
            step(Phase<2>(), k + 2);
            step(Phase<3>(), k + 3);
        }
        if (k <= k1) {
            step(Phase<0>(), k);
            if (k + 1 <= k1) {
                step(Phase<1>(), k + 1);
                if (k + 2 <= k1) step(Phase<2>(), k + 2);
            }
        }
    }
}

// ---- fused blend + collapse (+ output) -----------------------------------
template <class P> __device__ __forceinline__ P blend_px(double wc, P w, P b) {
    // CartesianBlendFunctor (src/blend.h:47-70), selected rather than branched
    const P mix = from_real<P>((wc * (double)w) + ((1.0 - wc) * (double)b));
    return wc >= 1.0 ? w : (wc <= 0.0 ? b : mix);
}

// SpecializedConvertPyramidToScalarFunctor with the deterministic dither
// draw (src/fixmath.h:2157-2294), without per-lane branches.
// h: the pixel's dither hash (dither_hash), shared by its channels.
template <class P, class I> struct OutConv {
    __device__ static __forceinline__ I conv(P v, int frac, bool det, uint32_t h, int c) {
        constexpr int32_t hi = sizeof(I) == 1 ? 255 : 65535;
        const int32_t q = (int32_t)v >> frac, f = (int32_t)v & ((1 << frac) - 1);
        const int32_t q1 = 1 << (frac - 2), q3 = 3 << (frac - 2);
        const bool up_rand = det ? dither_draw(h, c, frac) <= f : f >= (1 << (frac - 1));
        const bool band = f >= q1 && f < q3;
        const int32_t r = (band ? up_rand : f >= q1) ? q + 1 : q;
        return (I)min(hi, max(0, r));
    }
};
template <> struct OutConv<double, float> {
    __device__ static __forceinline__ float conv(double v, int, bool, uint32_t, int) {
        return (float)log_pyramid_to_float(v);
    }
};

// Output pixel pair stores into an interleaved row.  u0 / u1: write pixel X /
// X+1 (ownership, row end and the union alpha folded in).  A pair whose
// pixels are both written is one access where the row alignment allows;
// waves holding a pair with one pixel written (`split`, wave-uniform) add the
// per-pixel stores.
template <class I, int C, bool WIDE> struct OutPair {
    static constexpr bool RGB16 = WIDE && sizeof(I) == 2 && C == 3;
    static constexpr bool ONE = WIDE && C == 1 && sizeof(I) <= 2;
    __device__ static __forceinline__ void store(__amdgpu_buffer_rsrc_t rs, int X, const I (&v0)[C], const I (&v1)[C],
                                                 bool u0, bool u1, bool split) {
        if constexpr (RGB16) {
            const uint32_t o = 6u * (uint32_t)X;
            typedef uint32_t u3 __attribute__((ext_vector_type(3)));
            const uint32_t w0 = (uint32_t)v0[0] | ((uint32_t)v0[1] << 16);
            const uint32_t w1 = (uint32_t)v0[2] | ((uint32_t)v1[0] << 16);
            const uint32_t w2 = (uint32_t)v1[1] | ((uint32_t)v1[2] << 16);
            __builtin_amdgcn_raw_buffer_store_b96(u3{w0, w1, w2}, rs, u0 && u1 ? o : kBufSkip, 0, 0);
            if (split) {
                const uint32_t a0 = u0 && !u1 ? o : kBufSkip, a1 = u1 && !u0 ? o : kBufSkip;
                __builtin_amdgcn_raw_buffer_store_b32(w0, rs, a0, 0, 0);
                __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v0[2], rs, a0 + 4u, 0, 0);
                __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v1[0], rs, a1 + 6u, 0, 0);
                __builtin_amdgcn_raw_buffer_store_b32(w2, rs, a1 + 8u, 0, 0);
            }
        } else if constexpr (ONE && sizeof(I) == 2) {
            const uint32_t o = 2u * (uint32_t)X;
            __builtin_amdgcn_raw_buffer_store_b32((uint32_t)v0[0] | ((uint32_t)v1[0] << 16), rs, u0 && u1 ? o : kBufSkip, 0, 0);
            if (split) {
                __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v0[0], rs, u0 && !u1 ? o : kBufSkip, 0, 0);
                __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v1[0], rs, u1 && !u0 ? o + 2u : kBufSkip, 0, 0);
            }
        } else if constexpr (ONE) {
            const uint32_t o = (uint32_t)X;
            __builtin_amdgcn_raw_buffer_store_b16((unsigned short)((uint32_t)v0[0] | ((uint32_t)v1[0] << 8)), rs,
                                                  u0 && u1 ? o : kBufSkip, 0, 0);
            if (split) {
                __builtin_amdgcn_raw_buffer_store_b8((unsigned char)v0[0], rs, u0 && !u1 ? o : kBufSkip, 0, 0);
                __builtin_amdgcn_raw_buffer_store_b8((unsigned char)v1[0], rs, u1 && !u0 ? o + 1u : kBufSkip, 0, 0);
            }
        } else if constexpr (sizeof(I) == 4 && C == 3) {
            typedef uint32_t u3 __attribute__((ext_vector_type(3)));
            const uint32_t o = 12u * (uint32_t)X;
            __builtin_amdgcn_raw_buffer_store_b96(u3{__float_as_uint(v0[0]), __float_as_uint(v0[1]), __float_as_uint(v0[2])}, rs,
                                                  u0 ? o : kBufSkip, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b96(u3{__float_as_uint(v1[0]), __float_as_uint(v1[1]), __float_as_uint(v1[2])}, rs,
                                                  u1 ? o + 12u : kBufSkip, 0, 0);
        } else {
            constexpr uint32_t sz = sizeof(I);
            const uint32_t o0 = u0 ? (uint32_t)X * C * sz : kBufSkip, o1 = u1 ? ((uint32_t)X + 1u) * C * sz : kBufSkip;
#pragma unroll
            for (int ch = 0; ch < C; ++ch) {
                if constexpr (sz == 1) {
                    __builtin_amdgcn_raw_buffer_store_b8((unsigned char)v0[ch], rs, o0 + ch, 0, 0);
                    __builtin_amdgcn_raw_buffer_store_b8((unsigned char)v1[ch], rs, o1 + ch, 0, 0);
                } else if constexpr (sz == 2) {
                    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v0[ch], rs, o0 + 2u * ch, 0, 0);
                    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v1[ch], rs, o1 + 2u * ch, 0, 0);
                } else {
                    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v0[ch]), rs, o0 + 4u * ch, 0, 0);
                    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v1[ch]), rs, o1 + 4u * ch, 0, 0);
                }
            }
        }
    }
};

// One pyramid element of a plane row at column j (coarse level reads).
template <class P> __device__ __forceinline__ P bload_elem(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    if constexpr (sizeof(P) == 2) return (P)(int16_t)(uint16_t)bload_u16(r, off);
    else if constexpr (sizeof(P) == 4) return (P)bload<uint32_t>(r, off);
    else return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, bload<v2i32>(r, off)));
}

// Streaming structure: step i blends fine rows 2i, 2i+1 of the white and
// black Laplacian levels, adds the expansion of collapsed coarse rows i-1, i,
// i+1 and writes the rows (at level 0, converted into the output image with
// the union alpha).  Raw fine rows live in a 4-slot ring and raw coarse rows
// in a 2-slot ring; the loop is unrolled by their period of 2 steps, so no
// in-flight register is copied.  Step i loads fine rows 2i+2, 2i+3 and coarse
// row i+2 for step i+1 and issues the same memory instructions whatever it
// writes (zero-size resources and out-of-range offsets mask rows and lanes),
// so the waits for the previous step's loads count exactly.
//
// I = void: the collapsed level goes back into `black` (levels >= 1);
// otherwise it is converted into the interleaved output image of type I.
// WIDE: image rows 4-byte aligned and alpha / mask rows 2-byte aligned.
//
// OTF (level 0, integer pyramids): the white and black Laplacian values are
// formed on the fly from the images and their Gaussian levels 1, as the
// Laplacian pass would have stored them (v << frac minus the expansion,
// wrapping), so level 0 of neither Laplacian pyramid is written or read.
struct ImgArgs { const void* src; long spitch; int frac; };

template <class P, class M, int C, class I, bool MASK8, bool WIDE, bool OTF>
__global__ __launch_bounds__(256) void collapse_level_kernel(FusedCollapseArgs a)
{
    typedef Pix<P> Px;
    typedef typename Px::A A;
    typedef typename Px::S S;
    typedef BPair<P> B;
    constexpr bool OUT = !std::is_same<I, void>::value;
    typedef typename std::conditional<OUT, I, uint8_t>::type IT;
    typedef IntImageSrc<IT, P, C, WIDE> ISrc;
    static_assert(!OTF || (OUT && std::is_integral<P>::value), "on-the-fly Laplacians: level 0, integer pyramids");
    constexpr uint32_t ES = sizeof(P);
    // raw words of one fine row for this lane's pair
    struct LapPairs { typename B::Pair b[C], w[C]; };
    struct ImgPairs { typename ISrc::Raw b, w; };
    struct FineRaw {
        typename std::conditional<OTF, ImgPairs, LapPairs>::type v;
        typename std::conditional<MASK8, AlphaPair<WIDE>, typename BPair<M>::Pair>::type m;
        AlphaPair<WIDE> wa, ba;   // level 0 only
    };
    struct CoarseRaw { P v[C]; };

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
    LaneCols L;
    L.valid = owned;            // fine data is only read for owned columns
    L.both = xb;
    L.tailw = !wave_all(!owned || xb);
    L.c = X;
    const bool fw_even = (fw & 1) == 0;
    const int hdA = expand_hsum(X, cw, false, fw_even);
    const int hdB = expand_hsum(X + 1, cw, false, fw_even);
    const bool colp2 = wave_all(!owned || (((hdA & (hdA - 1)) == 0) && ((hdB & (hdB - 1)) == 0)));
    const int sA = __builtin_ctz((unsigned)hdA), sB = __builtin_ctz((unsigned)hdB);
    const bool cfirst = j == 0, clast = j == cw - 1 && j != 0;
    const uint32_t offc = valid ? (uint32_t)j * ES : kBufSkip;
    const uint32_t off2 = owned && xb ? (uint32_t)X * ES : kBufSkip;
    const uint32_t off1 = owned && !xb ? (uint32_t)X * ES : kBufSkip;
    const uint32_t fbytes = round_up2((uint32_t)fw) * ES;   // pyramid rows: pitch-padded pairs

    auto fetch_fine = [&](int Y, bool want, FineRaw& F) {
        const bool ok = want && Y < fh;
        const long y = ok ? Y : 0;
        const uint32_t nb = ok ? fbytes : 0u;
        const uint32_t off = owned ? (uint32_t)X * ES : kBufSkip;
        const P* brow = static_cast<const P*>(a.black) + y * a.fpitch;
        const P* wrow = static_cast<const P*>(a.white) + y * a.wpitch;
        if constexpr (OTF) {
            ISrc::fetch(ImgArgs{a.wimg, a.wipitch, a.ifrac}, fw, (int)y, ok, L, F.v.w);
            ISrc::fetch(ImgArgs{a.img, a.ipitch, a.ifrac}, fw, (int)y, ok, L, F.v.b);
        } else {
#pragma unroll
            for (int ch = 0; ch < C; ++ch) {
                F.v.b[ch] = B::fetch(buf_rsrc(brow + ch * a.fplane, nb), off);
                F.v.w[ch] = B::fetch(buf_rsrc(wrow + ch * a.wplane, nb), off);
            }
        }
        if constexpr (MASK8) {
            F.m.fetch(static_cast<const uint8_t*>(a.mask) + y * a.mpitch, ok, fw, L);
        } else {
            const M* mrow = static_cast<const M*>(a.mask) + y * a.mpitch;
            F.m = BPair<M>::fetch(buf_rsrc(mrow, ok ? round_up2((uint32_t)fw) * (uint32_t)sizeof(M) : 0u),
                                  owned ? (uint32_t)X * (uint32_t)sizeof(M) : kBufSkip);
        }
        if constexpr (OUT) {
            F.wa.fetch(a.walpha + y * a.wapitch, ok, fw, L);
            F.ba.fetch(a.balpha + y * a.bapitch, ok, fw, L);
        }
    };
    auto fetch_level = [&](const void* base, long pitch, long plane, int i, bool want, CoarseRaw& R) {
        const bool ok = want && i >= 0 && i < chh;
        const P* row = static_cast<const P*>(base) + (long)(ok ? i : 0) * pitch;
        const uint32_t nb = ok ? (uint32_t)cw * ES : 0u;
#pragma unroll
        for (int ch = 0; ch < C; ++ch) R.v[ch] = bload_elem<P>(buf_rsrc(row + ch * plane, nb), offc);
    };
    auto fetch_coarse = [&](int i, bool want, CoarseRaw& R) { fetch_level(a.coarse, a.cpitch, a.cplane, i, want, R); };
    // polyphase sums of a coarse row held one value per lane (no wraparound)
    auto coarse_sums = [&](const CoarseRaw& R, S (&hA)[C], S (&hB)[C]) {
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            const S x = Px::load(R.v[ch]);
            const S l = lane_up(x), r = lane_down(x);
            const S t = A::add(cfirst ? A::zero() : l, A::mul(x, 6));
            hA[ch] = clast ? t : A::add(t, r);
            hB[ch] = clast ? A::mul(x, 4) : A::mul(A::add(x, r), 4);
        }
    };
    // blend weight (double)m / white_value; white_value = 255 * 2^f takes
    // the exact quotient by 255 times 2^-f for 0 <= m < 2^24
    auto weight = [&](M m) -> double {
        if (a.wc_scale != 0.0 && wave_all((uint32_t)(int32_t)m < (1u << 24))) return quot255((int32_t)m) * a.wc_scale;
        return (double)m / a.white_value;
    };
    const bool det = a.dither == ENBL_DITHER_DET;

    // blend + add the expansion e (per channel) + store of fine row Y; with
    // OTF, ew / eb are the expansions of the white / black Gaussian levels 1
    auto finish = [&](int Y, const FineRaw& F, const S (&e0)[C], const S (&e1)[C], const S (&ew0)[C],
                      const S (&ew1)[C], const S (&eb0)[C], const S (&eb1)[C]) {
        const bool ok = Y < fh;
        M m0, m1;
        if constexpr (MASK8) {
            m0 = (M)((M)(F.m.w[0] & 0xffu) << a.mfrac);
            m1 = (M)((M)(WIDE ? (F.m.w[0] >> 8) & 0xffu : F.m.w[WIDE ? 0 : 1]) << a.mfrac);
        } else {
            typename Pix<M>::S t0, t1;
            Vec2<M>::split(F.m, t0, t1);
            m0 = (M)t0;
            m1 = (M)t1;
        }
        const double wc0 = weight(m0), wc1 = weight(m1);
        P o0[C], o1[C];
        S bi0[C], bi1[C], wi0[C], wi1[C];
        if constexpr (OTF) {
            ISrc::unpack(ImgArgs{a.img, a.ipitch, a.ifrac}, F.v.b, bi0, bi1);
            ISrc::unpack(ImgArgs{a.wimg, a.wipitch, a.ifrac}, F.v.w, wi0, wi1);
        }
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            P b0, b1, w0, w1;
            if constexpr (OTF) {
                b0 = Px::wrap(A::sub(bi0[ch], eb0[ch]));
                b1 = Px::wrap(A::sub(bi1[ch], eb1[ch]));
                w0 = Px::wrap(A::sub(wi0[ch], ew0[ch]));
                w1 = Px::wrap(A::sub(wi1[ch], ew1[ch]));
            } else {
                S t0, t1, u0, u1;
                Vec2<P>::split(F.v.b[ch], t0, t1);
                Vec2<P>::split(F.v.w[ch], u0, u1);
                b0 = (P)t0; b1 = (P)t1; w0 = (P)u0; w1 = (P)u1;
            }
            o0[ch] = Px::sat(A::add(Px::load(blend_px<P>(wc0, w0, b0)), e0[ch]));
            o1[ch] = Px::sat(A::add(Px::load(blend_px<P>(wc1, w1, b1)), e1[ch]));
        }
        if constexpr (!OUT) {
            const P* row = static_cast<const P*>(a.black) + (long)(ok ? Y : 0) * a.fpitch;
            const uint32_t nb = ok ? (uint32_t)fw * ES : 0u;
#pragma unroll
            for (int ch = 0; ch < C; ++ch) {
                const __amdgpu_buffer_rsrc_t rs = buf_rsrc(row + ch * a.fplane, nb);
                B::store2(rs, off2, o0[ch], o1[ch]);
                if (L.tailw) B::store1(rs, off1, o0[ch]);
            }
        } else {
            const bool wa0 = F.wa.first(), wa1 = F.wa.second();
            const bool ba0 = F.ba.first(), ba1 = F.ba.second();
            const bool u0 = owned && (wa0 || ba0);
            const bool u1 = owned && xb && (wa1 || ba1);
            IT v0[C], v1[C];
            const bool hashed = det && std::is_integral<P>::value;
            const uint32_t h0 = hashed ? dither_hash(a.seed, X, Y) : 0u, h1 = hashed ? dither_hash(a.seed, X + 1, Y) : 0u;
#pragma unroll
            for (int ch = 0; ch < C; ++ch) {
                v0[ch] = OutConv<P, IT>::conv(o0[ch], a.ifrac, det, h0, ch);
                v1[ch] = OutConv<P, IT>::conv(o1[ch], a.ifrac, det, h1, ch);
            }
            const long y = ok ? Y : 0;
            const IT* irow = static_cast<const IT*>(a.img) + y * a.ipitch;
            const uint32_t ib = ok ? (uint32_t)fw * C * (uint32_t)sizeof(IT) : 0u;
            const bool split = !wave_all(u0 == u1);
            OutPair<IT, C, WIDE>::store(buf_rsrc(irow, ib), X, v0, v1, u0, u1, split);
            // union alpha: black alpha = 255 where white alpha is set (src/enblend.h:461-463)
            const uint8_t* arow = a.balpha + y * a.bapitch;
            const uint32_t n0 = wa0 ? 255u : (WIDE ? F.ba.w[0] & 0xffu : F.ba.w[0]);
            const uint32_t n1 = wa1 ? 255u : (WIDE ? F.ba.w[0] >> 8 : F.ba.w[WIDE ? 0 : 1]);
            if constexpr (WIDE) {
                const __amdgpu_buffer_rsrc_t rs = buf_rsrc(arow, ok ? (uint32_t)fw : 0u);
                __builtin_amdgcn_raw_buffer_store_b16((unsigned short)(n0 | (n1 << 8)), rs, owned && xb ? (uint32_t)X : kBufSkip, 0, 0);
                if (L.tailw) __builtin_amdgcn_raw_buffer_store_b8((unsigned char)n0, rs, owned && !xb ? (uint32_t)X : kBufSkip, 0, 0);
            } else {
                const __amdgpu_buffer_rsrc_t rs = buf_rsrc(arow, ok ? (uint32_t)fw : 0u);
                __builtin_amdgcn_raw_buffer_store_b8((unsigned char)n0, rs, owned ? (uint32_t)X : kBufSkip, 0, 0);
                __builtin_amdgcn_raw_buffer_store_b8((unsigned char)n1, rs, owned && xb ? (uint32_t)X + 1u : kBufSkip, 0, 0);
            }
        }
    };

    // expansion of the coarse rows into fine rows 2i (e00/e10) and 2i+1 (e01/e11)
    auto expand2 = [&](int i, bool more, const S (&pA)[C], const S (&pB)[C], const S (&cA)[C], const S (&cB)[C],
                       const S (&nA)[C], const S (&nB)[C], S (&e00)[C], S (&e10)[C], S (&e01)[C], S (&e11)[C]) {
        const int Y = 2 * i;
        const int v0 = expand_vsum(Y, chh), v1 = expand_vsum(Y + 1, chh);
        const bool fast = std::is_same<A, AccI>::value && colp2 && (v0 & (v0 - 1)) == 0 && (v1 & (v1 - 1)) == 0;
        const int s0 = __builtin_ctz((unsigned)v0), s1 = __builtin_ctz((unsigned)v1);
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
            if constexpr (std::is_same<A, AccI>::value) {
                if (fast) {
                    e00[ch] = div_pow2(o00, sA + s0);
                    e10[ch] = div_pow2(o10, sB + s0);
                    e01[ch] = div_pow2(o01, sA + s1);
                    e11[ch] = div_pow2(o11, sB + s1);
                } else {
                    e00[ch] = o00 / (hdA * v0);
                    e10[ch] = o10 / (hdB * v0);
                    e01[ch] = o01 / (hdA * v1);
                    e11[ch] = o11 / (hdB * v1);
                }
            } else {
                e00[ch] = A::div(o00, hdA * v0);
                e10[ch] = A::div(o10, hdB * v0);
                e01[ch] = A::div(o01, hdA * v1);
                e11[ch] = A::div(o11, hdB * v1);
            }
        }
    };

    // polyphase sums of three consecutive rows (p, c) and the raw row ring of
    // one coarse level: the collapsed level l+1, and with OTF the white and
    // black Gaussian levels 1
    struct CoarseState {
        S pA[C], pB[C], cA[C], cB[C];
        CoarseRaw r[2];
    };
    constexpr int NS = OTF ? 3 : 1;
    CoarseState cs[NS];
    const void* cbase[3] = {a.coarse, a.wg1, a.bg1};
    const long cpitch[3] = {a.cpitch, a.wg1pitch, a.bg1pitch};
    const long cplane[3] = {a.cplane, a.wg1plane, a.bg1plane};
    FineRaw fr[4];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        CoarseRaw tp, tc;
        fetch_level(cbase[k], cpitch[k], cplane[k], i0 - 1, i0 > 0, tp);
        fetch_level(cbase[k], cpitch[k], cplane[k], i0, true, tc);
        fetch_level(cbase[k], cpitch[k], cplane[k], i0 + 1, true, cs[k].r[1]);
        coarse_sums(tp, cs[k].pA, cs[k].pB);
        if (i0 == 0) {
#pragma unroll
            for (int ch = 0; ch < C; ++ch) cs[k].pA[ch] = cs[k].pB[ch] = A::zero();
        }
        coarse_sums(tc, cs[k].cA, cs[k].cB);
    }
    // With OTF the fine rows are fetched in their own step (no prefetch):
    // the registers of a second pair of rows would cost a wave per SIMD.
    constexpr bool PREFETCH = !OTF;
    if constexpr (PREFETCH) {
        fetch_fine(2 * i0, true, fr[0]);
        fetch_fine(2 * i0 + 1, true, fr[1]);
    }

    auto step = [&](auto ph, int i) {
        constexpr int PH = decltype(ph)::value;
        constexpr int F0 = 2 * PH, F1 = 2 * PH + 1, N0 = 2 * (1 - PH), N1 = 2 * (1 - PH) + 1;
        const bool next = i + 1 < i1;
#pragma unroll
        for (int k = 0; k < NS; ++k) fetch_level(cbase[k], cpitch[k], cplane[k], i + 2, next, cs[k].r[PH]);
        if constexpr (PREFETCH) {
            fetch_fine(2 * i + 2, next, fr[N0]);
            fetch_fine(2 * i + 3, next, fr[N1]);
        } else {
            fetch_fine(2 * i, true, fr[F0]);
            fetch_fine(2 * i + 1, true, fr[F1]);
        }
        const bool more = i + 1 < chh;
        S e[NS][4][C];   // per level: rows 2i (pair A, B) and 2i+1 (pair A, B)
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            S nA[C], nB[C];
            coarse_sums(cs[k].r[1 - PH], nA, nB);
            expand2(i, more, cs[k].pA, cs[k].pB, cs[k].cA, cs[k].cB, nA, nB, e[k][0], e[k][1], e[k][2], e[k][3]);
#pragma unroll
            for (int ch = 0; ch < C; ++ch) {
                cs[k].pA[ch] = cs[k].cA[ch]; cs[k].pB[ch] = cs[k].cB[ch];
                cs[k].cA[ch] = nA[ch]; cs[k].cB[ch] = nB[ch];
            }
        }
        constexpr int KW = OTF ? 1 : 0, KB = OTF ? 2 : 0;
        finish(2 * i, fr[F0], e[0][0], e[0][1], e[KW][0], e[KW][1], e[KB][0], e[KB][1]);
        finish(2 * i + 1, fr[F1], e[0][2], e[0][3], e[KW][2], e[KW][3], e[KB][2], e[KB][3]);
    };

    int i = i0;
    for (; i + 1 < i1; i += 2) {
        step(Phase<0>(), i);
        step(Phase<1>(), i + 1);
    }
    if (i < i1) step(Phase<0>(), i);
}

// Strip shape of the fused passes: rows per strip chosen so that about
// g_strip_waves waves cover the level, at most g_strip_max_rows
// (enbl_set_strip_shape).  A strip also reads the halo rows of its coarse
// level above and below, so short strips read up to twice their rows: a
// strip has at least 4 rows, and up to 12 where the level still fills about
// 2048 waves (the fine level 1 of a 7168-column step: 104 -> 91 us for the
// two Laplacian passes).  Heights are rounded up to even: 11-row strips
// there measured 15% slower than 12-row ones, 10-row ones too.
int g_strip_waves = 8192, g_strip_max_rows = 32;

inline int strip_rows_f(int waves_x, int out_rows)
{
    const long cells = (long)waves_x * out_rows;
    long r = (cells + g_strip_waves - 1) / g_strip_waves;
    const long least = std::max(4L, std::min(12L, cells / 2048));
    if (r < least) r = least;
    r += r & 1;
    if (r > g_strip_max_rows) r = g_strip_max_rows;
    return (int)r;
}

template <class P, int C, class SRC, bool AW>
void launch_lap_t(FusedLapArgs a, hipStream_t s)
{
    const int waves_x = (a.cw + kOwnedF - 1) / kOwnedF + 1;
    if (a.rows <= 0) a.rows = strip_rows_f(waves_x, a.ch);
    dim3 grid((waves_x + 3) / 4, (a.ch + a.rows - 1) / a.rows);
    hipLaunchKernelGGL((laplacian_level_kernel<P, C, SRC, AW>), grid, dim3(256), 0, s, a);
}

// Alpha rows 2-byte aligned: one 16-bit load per pixel pair.
inline bool alpha_wide(const FusedLapArgs& a) { return ((uintptr_t)a.salpha & 1) == 0 && (a.sapitch & 1) == 0; }

template <class P, class SRC1, class SRC3>
void launch_lap_c(const FusedLapArgs& a, int C, hipStream_t s)
{
    const bool aw = alpha_wide(a);
    if (C == 1) { if (aw) launch_lap_t<P, 1, SRC1, true>(a, s); else launch_lap_t<P, 1, SRC1, false>(a, s); }
    else { if (aw) launch_lap_t<P, 3, SRC3, true>(a, s); else launch_lap_t<P, 3, SRC3, false>(a, s); }
}

template <class P, class M, class I, bool MASK8, bool WIDE>
void launch_col_t(FusedCollapseArgs a, int C, hipStream_t s)
{
    const int waves_x = (a.cw + kOwned - 1) / kOwned + 1;
    if (a.rows <= 0) a.rows = strip_rows_f(waves_x, a.ch);
    dim3 grid((waves_x + 3) / 4, (a.ch + a.rows - 1) / a.rows);
    constexpr bool CAN_OTF = !std::is_same<I, void>::value && std::is_integral<P>::value;
    if (CAN_OTF && a.otf) {
        if constexpr (CAN_OTF) {
            if (C == 1) hipLaunchKernelGGL((collapse_level_kernel<P, M, 1, I, MASK8, WIDE, true>), grid, dim3(256), 0, s, a);
            else hipLaunchKernelGGL((collapse_level_kernel<P, M, 3, I, MASK8, WIDE, true>), grid, dim3(256), 0, s, a);
        }
        return;
    }
    if (C == 1) hipLaunchKernelGGL((collapse_level_kernel<P, M, 1, I, MASK8, WIDE, false>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((collapse_level_kernel<P, M, 3, I, MASK8, WIDE, false>), grid, dim3(256), 0, s, a);
}

// Output rows aligned for whole-pair stores and 8-bit planes (mask, alphas)
// 2-byte aligned.
template <class I> bool collapse_wide(const FusedCollapseArgs& a, bool mask8)
{
    const auto even = [](const void* p, long pitch) { return ((uintptr_t)p & 1) == 0 && (pitch & 1) == 0; };
    if (mask8 && !even(a.mask, a.mpitch)) return false;
    if constexpr (!std::is_same<I, void>::value) {
        if (!even(a.walpha, a.wapitch) || !even(a.balpha, a.bapitch)) return false;
        if (sizeof(I) == 2 && (((uintptr_t)a.img & 3) != 0 || (a.ipitch & 1) != 0)) return false;
        if (sizeof(I) == 1 && !even(a.img, a.ipitch)) return false;
        // on-the-fly Laplacians read the white image with the same word loads
        if (a.otf && sizeof(I) == 2 && (((uintptr_t)a.wimg & 3) != 0 || (a.wipitch & 1) != 0)) return false;
        if (a.otf && sizeof(I) == 1 && !even(a.wimg, a.wipitch)) return false;
    }
    return true;
}

template <class P, class M, class I, bool MASK8>
void launch_col_w(const FusedCollapseArgs& a0, int C, hipStream_t s)
{
    FusedCollapseArgs a = a0;
    // white_value = 255 * 2^f: blend weights by the exact quotient by 255
    const double f = a.white_value / 255.0;
    a.wc_scale = (f >= 1.0 && f == ldexp(1.0, ilogb(f))) ? 1.0 / f : 0.0;
    if (collapse_wide<I>(a, MASK8)) launch_col_t<P, M, I, MASK8, true>(a, C, s);
    else launch_col_t<P, M, I, MASK8, false>(a, C, s);
}

}  // namespace

int launch_laplacian_level(int ptype, int src_kind, const FusedLapArgs& a, int C, hipStream_t s)
{
    if (C != 1 && C != 3) return ENBL_E_UNSUPPORTED;
    if (src_kind == 0) {
        if (ptype == ENBL_I16) { launch_lap_c<int16_t, PlaneSrc<int16_t, 1>, PlaneSrc<int16_t, 3>>(a, C, s); return 0; }
        if (ptype == ENBL_I32) { launch_lap_c<int32_t, PlaneSrc<int32_t, 1>, PlaneSrc<int32_t, 3>>(a, C, s); return 0; }
        if (ptype == ENBL_F64) { launch_lap_c<double, PlaneSrc<double, 1>, PlaneSrc<double, 3>>(a, C, s); return 0; }
    } else if (src_kind == ENBL_U8 && ptype == ENBL_I16) {
        // one 8-bit channel: pairs are 2-byte words when the rows are 2-byte aligned
        const bool iw = ((uintptr_t)a.src & 1) == 0 && (a.spitch & 1) == 0;
        if (C == 1 && iw) launch_lap_c<int16_t, IntImageSrc<uint8_t, int16_t, 1, true>, IntImageSrc<uint8_t, int16_t, 3, false>>(a, C, s);
        else launch_lap_c<int16_t, IntImageSrc<uint8_t, int16_t, 1, false>, IntImageSrc<uint8_t, int16_t, 3, false>>(a, C, s);
        return 0;
    } else if (src_kind == ENBL_U16 && ptype == ENBL_I32) {
        // 16-bit samples: pairs are whole 32-bit words when the rows are 4-byte aligned
        const bool iw = ((uintptr_t)a.src & 3) == 0 && (a.spitch & 1) == 0;
        if (iw) launch_lap_c<int32_t, IntImageSrc<uint16_t, int32_t, 1, true>, IntImageSrc<uint16_t, int32_t, 3, true>>(a, C, s);
        else launch_lap_c<int32_t, IntImageSrc<uint16_t, int32_t, 1, false>, IntImageSrc<uint16_t, int32_t, 3, false>>(a, C, s);
        return 0;
    } else if (src_kind == ENBL_F32 && ptype == ENBL_F64) {
        launch_lap_c<double, FloatImageSrc<1>, FloatImageSrc<3>>(a, C, s); return 0;
    }
    return ENBL_E_UNSUPPORTED;
}

int launch_collapse_level(int ptype, int mtype, int out_itype, bool mask8, const FusedCollapseArgs& a, int C, hipStream_t s)
{
    if (C != 1 && C != 3) return ENBL_E_UNSUPPORTED;
#define ENBL_COL(P, M)                                                                                      \
    do {                                                                                                    \
        if (out_itype == 0) { if (mask8) launch_col_w<P, M, void, true>(a, C, s); else launch_col_w<P, M, void, false>(a, C, s); return 0; } \
        if (out_itype == ENBL_U8 && std::is_integral<P>::value) { launch_col_w<P, M, uint8_t, true>(a, C, s); return 0; }       \
        if (out_itype == ENBL_U16 && std::is_integral<P>::value) { launch_col_w<P, M, uint16_t, true>(a, C, s); return 0; }     \
    } while (0)
    if (ptype == ENBL_I16 && mtype == ENBL_I16) ENBL_COL(int16_t, int16_t);
    if (ptype == ENBL_I32 && mtype == ENBL_I32) ENBL_COL(int32_t, int32_t);
    if (ptype == ENBL_F64 && mtype == ENBL_I32) {
        if (out_itype == 0) { if (mask8) launch_col_w<double, int32_t, void, true>(a, C, s); else launch_col_w<double, int32_t, void, false>(a, C, s); return 0; }
        if (out_itype == ENBL_F32) { launch_col_w<double, int32_t, float, true>(a, C, s); return 0; }
    }
#undef ENBL_COL
    return ENBL_E_UNSUPPORTED;
}

}  // namespace enbl

extern "C" int enbl_set_strip_shape(int target_waves, int max_rows)
{
    if (target_waves < 64 || max_rows < 4 || max_rows > 4096) return enbl::host::fail(ENBL_E_ARG, "strip shape %d %d", target_waves, max_rows);
    enbl::g_strip_waves = target_waves;
    enbl::g_strip_max_rows = max_rows;
    return ENBL_OK;
}
