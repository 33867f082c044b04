// Per-output bodies for pyramid levels too small for the streaming passes:
// one thread per destination sample (reduce, blend) or per source sample
// (expand, collapse), 64 columns x 4 rows per 256-thread block, with every
// load of a thread's window issued before any of it is combined.  The
// direct kernels (pyramid.hip) run one body per launch; the multi-job
// kernel (tail.hip) runs many jobs of the enblend step in one launch.
// Same boundary forms and operation order as the streaming kernels.
#pragma once

#include "enbl_device.h"
#include "enbl_internal.h"
#include "pyr_common.h"

namespace enbl {

// REDUCE with or without alpha (src/pyramid.h:166-573, :601-848).
template <class P, int C, bool ALPHA, bool WRAP>
__device__ __forceinline__ void reduce_small_body(const ReduceArgs& a, int bx, int by)
{
    typedef Pix<P> Px;
    typedef typename Px::A A;
    typedef typename Px::S S;
    const int j = bx * 64 + (threadIdx.x & 63);
    const int i = by * 4 + (threadIdx.x >> 6);
    if (j >= a.dw || i >= a.dh) return;
    const int sw = a.sw, sh = a.sh;
    const int c = 2 * j;
    const P* src = static_cast<const P*>(a.src);
    const int rcase = WRAP ? 0 : (c == sw - 1 ? 1 : (c == sw - 2 ? 2 : 0));
    int col[5];
    bool in[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int x = c - 2 + k;
        in[k] = WRAP || (x >= 0 && x < sw);
        col[k] = WRAP ? wrap_index(x, sw) : (x < 0 ? 0 : (x >= sw ? sw - 1 : x));
    }
    // Issue every load of the 5x5 window first (clamped indices are always
    // inside the level), then evaluate the literal boundary forms.
    const int r = 2 * i;
    int rowi[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int y = r - 2 + k;
        rowi[k] = y < 0 ? 0 : (y >= sh ? sh - 1 : y);
    }
    uint8_t al[5][5];
    S raw[5][5][C];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
#pragma unroll
        for (int t = 0; t < 5; ++t) {
            al[k][t] = ALPHA ? *gmem(a.salpha + (long)rowi[k] * a.sapitch + col[t]) : (uint8_t)1;
#pragma unroll
            for (int ch = 0; ch < C; ++ch)
                raw[k][t][ch] = Px::load(*gmem(src + ch * a.splane + (long)rowi[k] * a.spitch + col[t]));
        }
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
        for (int t = 0; t < 5; ++t) m[t] = ALPHA ? ((in[t] && al[k][t]) ? 1 : 0) : 1;
        if (ALPHA) {
            int hm = m[0] + 4 * m[1] + 6 * m[2];
            if (rcase != 1) hm += 4 * m[3];
            if (rcase == 0) hm += m[4];
            mrow[k] = hm;
        }
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            S v[5];
#pragma unroll
            for (int t = 0; t < 5; ++t) {
                if (ALPHA) v[t] = m[t] ? raw[k][t][ch] : A::zero();
                else v[t] = (WRAP || in[t] || t < 2) ? raw[k][t][ch] : A::zero();
            }
            const S a0 = A::add(v[0], A::mul(v[1], 4));
            S hv;
            if (rcase == 1) {
                hv = ALPHA ? A::add(a0, A::mul(v[2], 6)) : A::add(a0, A::mul(v[2], 11));
            } else if (rcase == 2) {
                const S d4 = A::mul(v[3], 4);
                hv = ALPHA ? A::add(A::add(a0, A::mul(v[2], 6)), d4)
                           : A::add(A::add(A::add(a0, A::mul(v[2], 6)), d4), A::div(d4, 4));
            } else {
                hv = A::add(A::add(A::add(a0, A::mul(v[2], 6)), A::mul(v[3], 4)), v[4]);
            }
            hrow[k][ch] = hv;
        }
    }
    int ap = 0;
    if (ALPHA) {
        const int q0 = r == 0 ? 0 : mrow[0] + 4 * mrow[1];
        if (r == sh - 1) ap = q0 + 6 * mrow[2];
        else if (r == sh - 2) ap = q0 + 6 * mrow[2] + 4 * mrow[3];
        else ap = q0 + 6 * mrow[2] + 4 * mrow[3] + mrow[4];
    }
    P* dst = static_cast<P*>(a.dst);
#pragma unroll
    for (int ch = 0; ch < C; ++ch) {
        S v;
        if (ALPHA) {
            const S a0 = r == 0 ? A::zero() : A::add(hrow[0][ch], A::mul(hrow[1][ch], 4));
            if (r == sh - 1) v = A::add(a0, A::mul(hrow[2][ch], 6));
            else if (r == sh - 2) v = A::add(A::add(a0, A::mul(hrow[2][ch], 6)), A::mul(hrow[3][ch], 4));
            else v = A::add(A::add(A::add(a0, A::mul(hrow[2][ch], 6)), A::mul(hrow[3][ch], 4)), hrow[4][ch]);
            *gmem(dst + ch * a.dplane + (long)i * a.dpitch + j) = ap ? Px::wrap(A::div(v, ap)) : (P)0;
        } else {
            const S a0 = r == 0 ? A::mul(hrow[2][ch], 5) : A::add(hrow[0][ch], A::mul(hrow[1][ch], 4));
            if (r == sh - 1) {
                v = A::add(a0, A::mul(hrow[2][ch], 11));
            } else if (r == sh - 2) {
                const S d4 = A::mul(hrow[3][ch], 4);
                v = A::add(A::add(A::add(a0, A::mul(hrow[2][ch], 6)), d4), A::div(d4, 4));
            } else {
                v = A::add(A::add(A::add(a0, A::mul(hrow[2][ch], 6)), A::mul(hrow[3][ch], 4)), hrow[4][ch]);
            }
            *gmem(dst + ch * a.dplane + (long)i * a.dpitch + j) = Px::wrap(A::div(v, 256));
        }
    }
    if (ALPHA && a.dalpha) *gmem(a.dalpha + (long)i * a.dapitch + j) = ap ? 255 : 0;
}

// Horizontal polyphase sums of one source row at column j for the even
// (hA) and odd (hB) destination columns 2j, 2j+1, with the reference's
// boundary forms (src/pyramid.h:1199-1411).  x = row[j]; L, R are the left
// and right neighbours as loaded (column 0 / sw-1 wrap around with WRAP;
// without it the value at an edge is unused).  y: source row (the
// 1-pixel-wide wraparound form differs on row 0).
template <class A, class S, bool WRAP>
__device__ __forceinline__ void expand_hsums(S x, S L, S R, int j, int sw, int y, bool dst_even, S& hA, S& hB)
{
    if (sw == 1) {
        S r1 = A::zero();
        if (WRAP) r1 = dst_even ? x : A::mul(x, 4);
        if (WRAP && y == 0) {
            hA = dst_even ? A::add(A::add(r1, A::mul(x, 6)), x) : A::add(A::add(r1, A::mul(x, 6)), A::mul(x, 4));
            hB = A::mul(A::add(x, x), 4);
        } else {
            hA = A::add(r1, A::mul(x, 6));
            hB = A::mul(x, 4);
        }
    } else if (j == 0) {
        S r1 = A::zero();
        if (WRAP) r1 = dst_even ? L : A::mul(L, 4);
        hA = A::add(A::add(r1, A::mul(x, 6)), R);
        hB = A::mul(A::add(x, R), 4);
    } else if (j == sw - 1) {
        if (WRAP) {
            hA = dst_even ? A::add(A::add(L, A::mul(x, 6)), R) : A::add(A::add(L, A::mul(x, 6)), A::mul(R, 4));
            hB = A::mul(A::add(x, R), 4);
        } else {
            hA = A::add(L, A::mul(x, 6));
            hB = A::mul(x, 4);
        }
    } else {
        hA = A::add(A::add(L, A::mul(x, 6)), R);
        hB = A::mul(A::add(x, R), 4);
    }
}

// The expansion of source rows i-1, i, i+1 at column j into the 2x2
// destination block (2j.., 2i..): e[q][ch], q = dx + 2 dy.  Loads: the 3x3
// neighbourhood, rows and columns clamped into the level.
template <class P, int C, bool WRAP>
__device__ __forceinline__ void expand_block(const P* src, long spitch, long splane, int sw, int sh, int dw, int j, int i,
                                             typename Pix<P>::S (&e)[4][C])
{
    typedef Pix<P> Px;
    typedef typename Px::A A;
    typedef typename Px::S S;
    const bool dst_even = (dw & 1) == 0;
    const int jl = j == 0 ? (WRAP ? sw - 1 : 0) : j - 1;
    const int jr = j == sw - 1 ? (WRAP ? 0 : j) : j + 1;
    S xv[3][3][C];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int y = i - 1 + k;
        const long ry = (long)(y < 0 ? 0 : (y >= sh ? sh - 1 : y)) * spitch;
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            const P* row = src + ch * splane + ry;
            xv[k][0][ch] = Px::load(*gmem(row + jl));
            xv[k][1][ch] = Px::load(*gmem(row + j));
            xv[k][2][ch] = Px::load(*gmem(row + jr));
        }
    }
    S hA[3][C], hB[3][C];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int y = i - 1 + k;
        const bool in = y >= 0 && y < sh;
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            S a, b;
            expand_hsums<A, S, WRAP>(xv[k][1][ch], xv[k][0][ch], xv[k][2][ch], j, sw, y, dst_even, a, b);
            hA[k][ch] = in ? a : A::zero();
            hB[k][ch] = in ? b : A::zero();
        }
    }
    const bool more = i + 1 < sh;
    const int X = 2 * j, Y = 2 * i;
    const int hdA = expand_hsum(X, sw, WRAP, dst_even), hdB = expand_hsum(X + 1, sw, WRAP, dst_even);
    const int v0 = expand_vsum(Y, sh), v1 = expand_vsum(Y + 1, sh);
#pragma unroll
    for (int ch = 0; ch < C; ++ch) {
        S o00, o10, o01, o11;
        if (more) {
            o00 = A::add(A::add(hA[0][ch], A::mul(hA[1][ch], 6)), hA[2][ch]);
            o10 = A::add(A::add(hB[0][ch], A::mul(hB[1][ch], 6)), hB[2][ch]);
            o01 = A::add(hA[1][ch], hA[2][ch]);
            o11 = A::add(hB[1][ch], hB[2][ch]);
        } else {
            o00 = A::add(hA[0][ch], A::mul(hA[1][ch], 6));
            o10 = A::add(hB[0][ch], A::mul(hB[1][ch], 6));
            o01 = hA[1][ch];
            o11 = hB[1][ch];
        }
        e[0][ch] = A::div(o00, hdA * v0);
        e[1][ch] = A::div(o10, hdB * v0);
        e[2][ch] = A::div(o01, hdA * v1);
        e[3][ch] = A::div(o11, hdB * v1);
    }
}

// EXPAND with add (collapse) or subtract (Laplacian), read-modify-write of
// the destination (src/pyramid.h:873-1457).  The destination samples are
// loaded with the source window.
template <class P, int C, bool ADD, bool WRAP>
__device__ __forceinline__ void expand_small_body(const ExpandArgs& a, int bx, int by)
{
    typedef Pix<P> Px;
    typedef typename Px::A A;
    typedef typename Px::S S;
    const int j = bx * 64 + (threadIdx.x & 63);
    const int i = by * 4 + (threadIdx.x >> 6);
    const int sw = a.sw, sh = a.sh, dw = a.dw, dh = a.dh;
    if (j >= sw || i >= sh) return;
    const int X = 2 * j, Y = 2 * i;
    const bool xb = X + 1 < dw, yb = Y + 1 < dh;
    P* dst = static_cast<P*>(a.dst);
    S d[4][C];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int x = (q & 1) && xb ? X + 1 : X, y = (q >> 1) && yb ? Y + 1 : Y;
#pragma unroll
        for (int ch = 0; ch < C; ++ch) d[q][ch] = Px::load(*gmem(dst + ch * a.dplane + (long)y * a.dpitch + x));
    }
    S e[4][C];
    expand_block<P, C, WRAP>(static_cast<const P*>(a.src), a.spitch, a.splane, sw, sh, dw, j, i, e);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int dx = q & 1, dy = q >> 1;
        if ((dx && !xb) || (dy && !yb)) continue;
#pragma unroll
        for (int ch = 0; ch < C; ++ch)
            *gmem(dst + ch * a.dplane + (long)(Y + dy) * a.dpitch + X + dx) =
                ADD ? Px::sat(A::add(d[q][ch], e[q][ch])) : Px::wrap(A::sub(d[q][ch], e[q][ch]));
    }
}

// CartesianBlendFunctor (src/blend.h:47-70), selected rather than branched.
template <class P> __device__ __forceinline__ P blend_value(double wc, P w, P b)
{
    const P mix = from_real<P>((wc * (double)w) + ((1.0 - wc) * (double)b));
    return wc >= 1.0 ? w : (wc <= 0.0 ? b : mix);
}

// One level of the blend (src/blend.h:79-121): black = blend(mask, white, black).
struct BlendArgs {
    const void* mask; long mpitch;
    const void* white; long wpitch; long wplane;
    void* black; long bpitch; long bplane;
    int w, h;
    double white_value;
};

template <class P, class M, int C>
__device__ __forceinline__ void blend_small_body(const BlendArgs& a, int bx, int by)
{
    const int x = bx * 64 + (threadIdx.x & 63);
    const int y = by * 4 + (threadIdx.x >> 6);
    if (x >= a.w || y >= a.h) return;
    const double wc = (double)*gmem(static_cast<const M*>(a.mask) + (long)y * a.mpitch + x) / a.white_value;
    const P* wh = static_cast<const P*>(a.white) + (long)y * a.wpitch + x;
    P* bk = static_cast<P*>(a.black) + (long)y * a.bpitch + x;
    P wv[C], bv[C];
#pragma unroll
    for (int ch = 0; ch < C; ++ch) {
        wv[ch] = *gmem(wh + ch * a.wplane);
        bv[ch] = *gmem(bk + ch * a.bplane);
    }
#pragma unroll
    for (int ch = 0; ch < C; ++ch) *gmem(bk + ch * a.bplane) = blend_value<P>(wc, wv[ch], bv[ch]);
}

// One collapse level (src/pyramid.h:1757-1784) with its blend fused in:
// black_l = sat(blend(mask_l, white_l, black_l) + expand(collapsed_{l+1})),
// no wraparound.  One thread per coarse sample; `coarse` is the collapsed
// level l+1, `fine` the level-l planes.
struct CollapseSmallArgs {
    const void* coarse; long cpitch; long cplane; int cw, ch;
    BlendArgs fine;
};

template <class P, class M, int C>
__device__ __forceinline__ void collapse_small_body(const CollapseSmallArgs& a, int bx, int by)
{
    typedef Pix<P> Px;
    typedef typename Px::A A;
    typedef typename Px::S S;
    const int j = bx * 64 + (threadIdx.x & 63);
    const int i = by * 4 + (threadIdx.x >> 6);
    const BlendArgs& f = a.fine;
    if (j >= a.cw || i >= a.ch) return;
    const int X = 2 * j, Y = 2 * i;
    const bool xb = X + 1 < f.w, yb = Y + 1 < f.h;
    P wv[4][C], bv[4][C];
    M mv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int x = (q & 1) && xb ? X + 1 : X, y = (q >> 1) && yb ? Y + 1 : Y;
        mv[q] = *gmem(static_cast<const M*>(f.mask) + (long)y * f.mpitch + x);
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            wv[q][ch] = *gmem(static_cast<const P*>(f.white) + ch * f.wplane + (long)y * f.wpitch + x);
            bv[q][ch] = *gmem(static_cast<const P*>(f.black) + ch * f.bplane + (long)y * f.bpitch + x);
        }
    }
    S e[4][C];
    expand_block<P, C, false>(static_cast<const P*>(a.coarse), a.cpitch, a.cplane, a.cw, a.ch, f.w, j, i, e);
    P* bk = static_cast<P*>(f.black);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int dx = q & 1, dy = q >> 1;
        if ((dx && !xb) || (dy && !yb)) continue;
        const double wc = (double)mv[q] / f.white_value;
#pragma unroll
        for (int ch = 0; ch < C; ++ch)
            *gmem(bk + ch * f.bplane + (long)(Y + dy) * f.bpitch + X + dx) =
                Px::sat(A::add(Px::load(blend_value<P>(wc, wv[q][ch], bv[q][ch])), e[q][ch]));
    }
}

}  // namespace enbl
