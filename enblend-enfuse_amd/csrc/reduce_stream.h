// Streaming reduce of one strip of a level (shared by the per-level launches
// of pyramid.hip and the batched passes of batch.hip).
//
// Reference: src/pyramid.h:166-573 (reduce with alpha), :601-848 (reduce
// without alpha).  Each 64-lane wave owns a strip of 62 output columns and
// streams down the rows of its strip:
//   * every lane loads the source samples of its own column pair with one
//     coalesced vector load per row and channel, and takes the horizontal
//     neighbours from the adjacent lanes through cross-lane shuffles;
//   * the horizontal 5-tap row sums are formed in registers and a rolling
//     window of row sums gives the vertical pass, so every source row is read
//     once per strip; the next output row's source rows are loaded before the
//     current one is computed;
//   * the boundary lanes evaluate the reference's literal boundary
//     expressions (clamp, zero-extension, wraparound and its 4x tap weights).
// The operation order of every sum matches the reference's SKIPSM state
// updates, which makes the double path bit-identical to it and the integer
// path exact.
#pragma once

#include "enbl_device.h"
#include "enbl_internal.h"
#include "pyr_common.h"
#include "row_fetch.h"

namespace enbl {

// Row sources of a reduce: a planar level, or (level 0 of an image pyramid)
// the interleaved integer image converted on the fly by
// ConvertScalarToPyramidFunctor, (P)v << frac (src/fixmath.h:116-163).
// `a` is any descriptor with the fields src, spitch, splane and frac.
template <class P> struct PlaneRows {
    static constexpr bool planar = true;
    typedef typename Pix<P>::S S;
    template <int C, class Args>
    __device__ static __forceinline__ void load(const Args& a, int r, int c, bool both, S (&x0)[C], S (&x1)[C]) {
        const P* src = static_cast<const P*>(a.src);
#pragma unroll
        for (int ch = 0; ch < C; ++ch) Vec2<P>::load(src + ch * a.splane + (long)r * a.spitch + c, both, x0[ch], x1[ch]);
    }
};

template <class I, class P> struct ImageRows {
    static constexpr bool planar = false;
    typedef typename Pix<P>::S S;
    template <int C, class Args>
    __device__ static __forceinline__ void load(const Args& a, int r, int c, bool both, S (&x0)[C], S (&x1)[C]) {
        const I* p = static_cast<const I*>(a.src) + (long)r * a.spitch + (long)c * C;
        const auto* g = gmem(p);
        I v[2 * C];
        if constexpr (sizeof(I) == 2 && C == 3) {
            // two RGB16 pixels = 12 bytes: one dwordx3 access when aligned
            if (both && ((uintptr_t)p & 3) == 0) {
                const auto* d = gmem(reinterpret_cast<const uint32_t*>(p));
                const uint32_t d0 = d[0], d1 = d[1], d2 = d[2];
                v[0] = (I)(d0 & 0xffffu); v[1] = (I)(d0 >> 16);
                v[2] = (I)(d1 & 0xffffu); v[3] = (I)(d1 >> 16);
                v[4] = (I)(d2 & 0xffffu); v[5] = (I)(d2 >> 16);
            } else {
#pragma unroll
                for (int k = 0; k < 2 * C; ++k) v[k] = (k < C || both) ? g[k] : (I)0;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 2 * C; ++k) v[k] = (k < C || both) ? g[k] : (I)0;
        }
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            x0[ch] = (S)(P)((P)v[ch] << a.frac);
            x1[ch] = both ? (S)(P)((P)v[C + ch] << a.frac) : (S)0;
        }
    }
};

// Raw samples of one source row held by one lane: its own pair (2j, 2j+1),
// alpha-masked, and, for wraparound edge lanes, the four taps that wrap to
// the far side of the row.
template <class S, int C, bool WRAP>
struct ReduceRow {
    S x0[C], x1[C];
    int m0, m1;
    S wa[WRAP ? C : 1], wb[WRAP ? C : 1], wd[WRAP ? C : 1], we[WRAP ? C : 1];
    int qa, qb, qd, qe;
};

template <class P, int C, bool ALPHA, bool WRAP, class SRC>
struct ReduceLane {
    static_assert(!WRAP || SRC::planar, "wraparound edge taps read planar levels");
    typedef Pix<P> Px;
    typedef typename Px::A A;
    typedef typename Px::S S;
    typedef ReduceRow<S, C, WRAP> Row;

    const ReduceArgs& a;
    int c;          // source centre column 2j
    bool valid;     // lane's destination column exists
    bool both;      // c + 1 is inside the row
    bool edge;      // a tap of this lane leaves the row

    // Issue the loads of source row r (no waiting, no cross-lane traffic).
    __device__ __forceinline__ void load(int r, Row& R) const {
        const P* src = static_cast<const P*>(a.src);
        R.m0 = 0;
        R.m1 = 0;
        if (ALPHA && valid) {
            const auto* ar = gmem(a.salpha + (long)r * a.sapitch);
            R.m0 = ar[c] ? 1 : 0;
            R.m1 = (both && ar[c + 1]) ? 1 : 0;
        }
#pragma unroll
        for (int ch = 0; ch < C; ++ch) { R.x0[ch] = A::zero(); R.x1[ch] = A::zero(); }
        if (valid) SRC::template load<C>(a, r, c, both, R.x0, R.x1);
        if (ALPHA) {
#pragma unroll
            for (int ch = 0; ch < C; ++ch) {
                if (!R.m0) R.x0[ch] = A::zero();
                if (!R.m1) R.x1[ch] = A::zero();
            }
        }
        if (WRAP) {
            R.qa = R.qb = R.qd = R.qe = 0;
#pragma unroll
            for (int ch = 0; ch < C; ++ch) { R.wa[ch] = R.wb[ch] = R.wd[ch] = R.we[ch] = A::zero(); }
            if (edge) {
                const int sw = a.sw;
                const int ia = wrap_index(c - 2, sw), ib = wrap_index(c - 1, sw);
                const int id = wrap_index(c + 1, sw), ie = wrap_index(c + 2, sw);
                int ma = 1, mb = 1, md = 1, me = 1;
                if (ALPHA) {
                    const uint8_t* ar = a.salpha + (long)r * a.sapitch;
                    ma = ar[ia] ? 1 : 0; mb = ar[ib] ? 1 : 0; md = ar[id] ? 1 : 0; me = ar[ie] ? 1 : 0;
                }
                R.qa = ma; R.qb = mb; R.qd = md; R.qe = me;
#pragma unroll
                for (int ch = 0; ch < C; ++ch) {
                    const P* row = src + ch * a.splane + (long)r * a.spitch;
                    R.wa[ch] = ma ? Px::load(row[ia]) : A::zero();
                    R.wb[ch] = mb ? Px::load(row[ib]) : A::zero();
                    R.wd[ch] = md ? Px::load(row[id]) : A::zero();
                    R.we[ch] = me ? Px::load(row[ie]) : A::zero();
                }
            }
        }
    }

    // Horizontal 5-tap sums of a loaded row (every lane of the wave executes
    // this: the neighbour taps arrive through shuffles).
    __device__ __forceinline__ void sums(const Row& R, S (&h)[C], int& hm) const {
        const int sw = a.sw;
        // boundary case of the reference's row tail (src/pyramid.h:264-299,
        // :410-455, :688-707): 1 = centre is the last sample, 2 = centre is
        // the second to last sample (no wraparound only)
        const int rcase = WRAP ? 0 : (c == sw - 1 ? 1 : (c == sw - 2 ? 2 : 0));
        if (ALPHA) {
            const int ml0 = lane_up(R.m0), ml1 = lane_up(R.m1), mr0 = lane_down(R.m0);
            int qa, qb, qd, qe;
            if (WRAP && edge) { qa = R.qa; qb = R.qb; qd = R.qd; qe = R.qe; }
            else { qa = c >= 2 ? ml0 : 0; qb = c >= 1 ? ml1 : 0; qd = R.m1; qe = mr0; }
            hm = qa + 4 * qb + 6 * R.m0;
            if (rcase != 1) hm += 4 * qd;
            if (rcase == 0) hm += qe;
        }
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            const S x0 = R.x0[ch];
            const S l0 = lane_up(x0), l1 = lane_up(R.x1[ch]), r0 = lane_down(x0);
            S va, vb, vd, ve;
            if (WRAP && edge) {
                va = R.wa[ch]; vb = R.wb[ch]; vd = R.wd[ch]; ve = R.we[ch];
            } else {
                va = c >= 2 ? l0 : (ALPHA ? A::zero() : x0);
                vb = c >= 1 ? l1 : (ALPHA ? A::zero() : x0);
                vd = R.x1[ch];
                ve = r0;
            }
            const S a0 = A::add(va, A::mul(vb, 4));
            if (rcase == 1) {
                h[ch] = ALPHA ? A::add(a0, A::mul(x0, 6)) : A::add(a0, A::mul(x0, 11));
            } else if (rcase == 2) {
                const S d4 = A::mul(vd, 4);
                h[ch] = ALPHA ? A::add(A::add(a0, A::mul(x0, 6)), d4)
                              : A::add(A::add(A::add(a0, A::mul(x0, 6)), d4), A::div(d4, 4));
            } else {
                h[ch] = A::add(A::add(A::add(a0, A::mul(x0, 6)), A::mul(vd, 4)), ve);
            }
        }
    }
};

// One wave's strip: output columns [62 wx - 1, 62 wx + 61) (lanes 1..62
// own them), output rows [i0, i0 + a.rows).
template <class P, int C, bool ALPHA, bool WRAP, class SRC>
__device__ __forceinline__ void reduce_strip(const ReduceArgs& a, int wx, int i0)
{
    typedef ReduceLane<P, C, ALPHA, WRAP, SRC> Lane;
    typedef typename Lane::A A;
    typedef typename Lane::S S;
    typedef typename Lane::Row Row;
    typedef Pix<P> Px;

    const int lane = threadIdx.x & (kWave - 1);
    if (wx * kOwned - 1 >= a.dw) return;  // whole wave beyond the level
    if (i0 >= a.dh) return;
    const int i1 = min(a.dh, i0 + a.rows);
    const int sh = a.sh;
    const int j = wx * kOwned + lane - 1;
    const int c = 2 * j;
    const bool valid = j >= 0 && j < a.dw;
    const bool owned = lane >= 1 && lane <= kOwned && valid;
    const Lane L{a, c, valid, c + 1 < a.sw, valid && (j == 0 || c + 2 >= a.sw)};
    P* dst = static_cast<P*>(a.dst);

    S hm2[C], hm1[C], hc[C], hp1[C], hp2[C];
    int qm2 = 0, qm1 = 0, qc = 0, qp1 = 0, qp2 = 0;
#pragma unroll
    for (int ch = 0; ch < C; ++ch) { hm2[ch] = hm1[ch] = hc[ch] = hp1[ch] = hp2[ch] = A::zero(); }
    Row ra, rb, na, nb;
    {
        Row t;
        if (i0 > 0) {
            L.load(2 * i0 - 2, t);
            L.sums(t, hm2, qm2);
            L.load(2 * i0 - 1, t);
            L.sums(t, hm1, qm1);
        }
        L.load(2 * i0, t);
        L.sums(t, hc, qc);
    }
    if (2 * i0 + 1 < sh) L.load(2 * i0 + 1, ra);
    if (2 * i0 + 2 < sh) L.load(2 * i0 + 2, rb);

    for (int i = i0; i < i1; ++i) {
        const int r = 2 * i;
        // prefetch the next output row's two source rows
        if (i + 1 < i1) {
            if (r + 3 < sh) L.load(r + 3, na);
            if (r + 4 < sh) L.load(r + 4, nb);
        }
        if (r + 1 < sh) L.sums(ra, hp1, qp1);
        if (r + 2 < sh) L.sums(rb, hp2, qp2);

        int ap = 0;
        if (ALPHA) {
            const int q0 = r == 0 ? 0 : qm2 + 4 * qm1;
            if (r == sh - 1) ap = q0 + 6 * qc;
            else if (r == sh - 2) ap = q0 + 6 * qc + 4 * qp1;
            else ap = q0 + 6 * qc + 4 * qp1 + qp2;
        }
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            S v;
            if (ALPHA) {
                // src/pyramid.h:350-384, :459-504 (column recursion)
                const S a0 = r == 0 ? A::zero() : A::add(hm2[ch], A::mul(hm1[ch], 4));
                if (r == sh - 1) v = A::add(a0, A::mul(hc[ch], 6));
                else if (r == sh - 2) v = A::add(A::add(a0, A::mul(hc[ch], 6)), A::mul(hp1[ch], 4));
                else v = A::add(A::add(A::add(a0, A::mul(hc[ch], 6)), A::mul(hp1[ch], 4)), hp2[ch]);
            } else {
                // src/pyramid.h:669-672, :741-760, :809-846 (clamped columns)
                const S a0 = r == 0 ? A::mul(hc[ch], 5) : A::add(hm2[ch], A::mul(hm1[ch], 4));
                if (r == sh - 1) {
                    v = A::add(a0, A::mul(hc[ch], 11));
                } else if (r == sh - 2) {
                    const S d4 = A::mul(hp1[ch], 4);
                    v = A::add(A::add(A::add(a0, A::mul(hc[ch], 6)), d4), A::div(d4, 4));
                } else {
                    v = A::add(A::add(A::add(a0, A::mul(hc[ch], 6)), A::mul(hp1[ch], 4)), hp2[ch]);
                }
            }
            if (owned) {
                P out;
                if (ALPHA) out = ap ? Px::wrap(A::div(v, ap)) : (P)0;
                else out = Px::wrap(A::div(v, 256));
                *gmem(dst + ch * a.dplane + (long)i * a.dpitch + j) = out;
            }
            hm2[ch] = hc[ch];
            hm1[ch] = hp1[ch];
            hc[ch] = hp2[ch];
        }
        if (ALPHA && owned && a.dalpha) *gmem(a.dalpha + (long)i * a.dapitch + j) = (uint8_t)(ap ? 255 : 0);
        qm2 = qc; qm1 = qp1; qc = qp2;
        ra = na;
        rb = nb;
    }
}

// The same strip without wraparound over raw-row rings: each step loads the
// two source rows of the next output row through per-row buffer resources
// (zero-size for rows outside the level or past the strip, out-of-range
// offsets for lanes outside it) into one half of a 4-slot ring and forms
// the row sums of the rows loaded by the previous step from the other half;
// the loop is unrolled by the ring's period of 2 steps, no in-flight
// register is copied, and every step issues the same memory instructions,
// so its wait covers the previous step's loads only.  SRCF: a row source of
// row_fetch.h; AW: alpha rows 2-byte aligned (one 16-bit load per pair).
// W0: also write the source rows as converted (the source's level 0, planar
// doubles at a.conv), each fine row by the strip that owns it: rows
// [2 i0, 2 i1) of the strip's coarse rows [i0, i1), columns by owned lanes.
template <class P, int C, bool ALPHA, class SRCF, bool AW, bool W0 = false>
__device__ __forceinline__ void reduce_strip_ring(const ReduceArgs& a, int wx, int i0)
{
    typedef Pix<P> Px;
    typedef typename Px::A A;
    typedef typename Px::S S;
    constexpr bool INT = std::is_same<A, AccI>::value;
    constexpr uint32_t ES = sizeof(P);
    struct RawRow {
        typename SRCF::Raw px;
        AlphaPair<AW> al;
    };

    const int lane = threadIdx.x & (kWave - 1);
    if (wx * kOwned - 1 >= a.dw) return;  // whole wave beyond the level
    if (i0 >= a.dh) return;
    const int i1 = min(a.dh, i0 + a.rows);
    const int sw = a.sw, sh = a.sh;
    const int j = wx * kOwned + lane - 1;
    const int c = 2 * j;
    const bool valid = j >= 0 && j < a.dw;
    const bool owned = lane >= 1 && lane <= kOwned && valid;
    const bool both = c + 1 < sw;
    // boundary case of the row tail (src/pyramid.h:264-299, :410-455,
    // :688-707): 1 = centre is the last sample, 2 = the second to last
    const int rcase = c == sw - 1 ? 1 : (c == sw - 2 ? 2 : 0);
    LaneCols L;
    L.valid = valid;
    L.both = both;
    L.tailw = !wave_all(!valid || both);
    L.c = c;

    auto fetch = [&](int r, bool want, RawRow& R) {
        const bool ok = want && r >= 0 && r < sh;
        const int rr = ok ? r : 0;
        SRCF::fetch(a, sw, rr, ok, L, R.px);
        if constexpr (ALPHA) R.al.fetch(a.salpha + (long)rr * a.sapitch, ok, sw, L);
    };
    // Horizontal 5-tap sums of a fetched row; the row-tail cases are
    // selected per lane, in the reference's operation order.
    auto sums = [&](const RawRow& R, S (&h)[C], int& hm, int wrow) {
        S x0[C], x1[C];
        SRCF::unpack(a, R.px, x0, x1);
        if constexpr (W0) {
            if (wrow >= 0 && owned) {
                const S* crow = static_cast<const S*>(a.conv) + (long)wrow * a.cpitch;
#pragma unroll
                for (int ch = 0; ch < C; ++ch) {
                    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(crow + ch * a.cplane, (uint32_t)sw * (uint32_t)sizeof(S));
                    if (both) BPair<S>::store2(rs, (uint32_t)c * (uint32_t)sizeof(S), x0[ch], x1[ch]);
                    else BPair<S>::store1(rs, (uint32_t)c * (uint32_t)sizeof(S), x0[ch]);
                }
            }
        } else {
            (void)wrow;
        }
        int m0 = 1, m1 = 1;
        if constexpr (ALPHA) {
            m0 = valid && R.al.first() ? 1 : 0;
            m1 = valid && both && R.al.second() ? 1 : 0;
            const int ml0 = lane_up(m0), ml1 = lane_up(m1), mr0 = lane_down(m0);
            const int qa = c >= 2 ? ml0 : 0, qb = c >= 1 ? ml1 : 0;
            hm = qa + 4 * qb + 6 * m0 + (rcase != 1 ? 4 * m1 : 0) + (rcase == 0 ? mr0 : 0);
        }
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            const S xa = (ALPHA ? m0 : valid) ? x0[ch] : A::zero();
            const S xb = (ALPHA ? m1 : (valid && both)) ? x1[ch] : A::zero();
            const S l0 = lane_up(xa), l1 = lane_up(xb), r0 = lane_down(xa);
            const S va = c >= 2 ? l0 : (ALPHA ? A::zero() : xa);
            const S vb = c >= 1 ? l1 : (ALPHA ? A::zero() : xa);
            const S a0 = A::add(va, A::mul(vb, 4));
            const S a6 = A::add(a0, A::mul(xa, 6));
            const S t0 = A::add(A::add(a6, A::mul(xb, 4)), r0);
            S t1, t2;
            if constexpr (ALPHA) {
                t1 = a6;
                t2 = A::add(a6, A::mul(xb, 4));
            } else {
                t1 = A::add(a0, A::mul(xa, 11));
                const S d4 = A::mul(xb, 4);
                t2 = A::add(A::add(a6, d4), A::div(d4, 4));
            }
            h[ch] = rcase == 1 ? t1 : (rcase == 2 ? t2 : t0);
        }
    };

    S hm2[C], hm1[C], hc[C], hp1[C], hp2[C];
    int qm2 = 0, qm1 = 0, qc = 0, qp1 = 0, qp2 = 0;
#pragma unroll
    for (int ch = 0; ch < C; ++ch) { hm2[ch] = hm1[ch] = hc[ch] = hp1[ch] = hp2[ch] = A::zero(); }
    RawRow ring[4];
    {
        RawRow t0, t1, t2;
        fetch(2 * i0 - 2, i0 > 0, t0);
        fetch(2 * i0 - 1, i0 > 0, t1);
        fetch(2 * i0, true, t2);
        fetch(2 * i0 + 1, true, ring[0]);
        fetch(2 * i0 + 2, true, ring[1]);
        if (i0 > 0) {
            sums(t0, hm2, qm2, -1);
            sums(t1, hm1, qm1, -1);
        }
        sums(t2, hc, qc, 2 * i0);
    }
    P* dst = static_cast<P*>(a.dst);
    const uint32_t offd = owned ? (uint32_t)j * ES : kBufSkip;
    const uint32_t offa = owned ? (uint32_t)j : kBufSkip;

    auto step = [&](auto ph, int i) {
        constexpr int PH = decltype(ph)::value;
        constexpr int F0 = 2 * PH, F1 = 2 * PH + 1, N0 = 2 * (1 - PH), N1 = 2 * (1 - PH) + 1;
        const int r = 2 * i;
        const bool next = i + 1 < i1;
        fetch(r + 3, next, ring[N0]);
        fetch(r + 4, next, ring[N1]);
        if (r + 1 < sh) sums(ring[F0], hp1, qp1, r + 1);
        if (r + 2 < sh) sums(ring[F1], hp2, qp2, next ? r + 2 : -1);

        int ap = 0;
        if (ALPHA) {
            const int q0 = r == 0 ? 0 : qm2 + 4 * qm1;
            if (r == sh - 1) ap = q0 + 6 * qc;
            else if (r == sh - 2) ap = q0 + 6 * qc + 4 * qp1;
            else ap = q0 + 6 * qc + 4 * qp1 + qp2;
        }
        S v[C];
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            if (ALPHA) {
                // src/pyramid.h:350-384, :459-504 (column recursion)
                const S a0 = r == 0 ? A::zero() : A::add(hm2[ch], A::mul(hm1[ch], 4));
                if (r == sh - 1) v[ch] = A::add(a0, A::mul(hc[ch], 6));
                else if (r == sh - 2) v[ch] = A::add(A::add(a0, A::mul(hc[ch], 6)), A::mul(hp1[ch], 4));
                else v[ch] = A::add(A::add(A::add(a0, A::mul(hc[ch], 6)), A::mul(hp1[ch], 4)), hp2[ch]);
            } else {
                // src/pyramid.h:669-672, :741-760, :809-846 (clamped columns)
                const S a0 = r == 0 ? A::mul(hc[ch], 5) : A::add(hm2[ch], A::mul(hm1[ch], 4));
                if (r == sh - 1) {
                    v[ch] = A::add(a0, A::mul(hc[ch], 11));
                } else if (r == sh - 2) {
                    const S d4 = A::mul(hp1[ch], 4);
                    v[ch] = A::add(A::add(A::add(a0, A::mul(hc[ch], 6)), d4), A::div(d4, 4));
                } else {
                    v[ch] = A::add(A::add(A::add(a0, A::mul(hc[ch], 6)), A::mul(hp1[ch], 4)), hp2[ch]);
                }
            }
        }
        P out[C];
        if constexpr (ALPHA && INT) {
            // power-of-two weight sums (full windows) divide by a shift, chosen per wave
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
            for (int ch = 0; ch < C; ++ch) out[ch] = ALPHA ? (ap ? Px::wrap(A::div(v[ch], ap)) : (P)0) : Px::wrap(A::div(v[ch], 256));
        }
        const P* drow = dst + (long)i * a.dpitch;
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            const __amdgpu_buffer_rsrc_t rs = buf_rsrc(drow + ch * a.dplane, (uint32_t)a.dw * ES);
            if constexpr (ES == 2) __builtin_amdgcn_raw_buffer_store_b16((unsigned short)out[ch], rs, offd, 0, 0);
            else if constexpr (ES == 4) __builtin_amdgcn_raw_buffer_store_b32((uint32_t)out[ch], rs, offd, 0, 0);
            else BPair<P>::store1(rs, offd, out[ch]);
        }
        if constexpr (ALPHA) {
            __builtin_amdgcn_raw_buffer_store_b8((unsigned char)(ap ? 255 : 0),
                                                 buf_rsrc(a.dalpha + (long)i * a.dapitch, a.dalpha ? (uint32_t)a.dw : 0u),
                                                 offa, 0, 0);
        }
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            hm2[ch] = hc[ch];
            hm1[ch] = hp1[ch];
            hc[ch] = hp2[ch];
        }
        qm2 = qc; qm1 = qp1; qc = qp2;
    };

    int i = i0;
    for (; i + 1 < i1; i += 2) {
        step(Phase<0>(), i);
        step(Phase<1>(), i + 1);
    }
    if (i < i1) step(Phase<0>(), i);
}


}  // namespace enbl
