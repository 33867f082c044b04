// Shared templates of the oracle (the CPU restatement of the reference's
// pyramid path).  TEST INFRASTRUCTURE ONLY -- see oracle_pyramid.cpp.
#pragma once

#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {


// ---------------------------------------------------------------------------
// Accumulator arithmetic with the reference's semantics.
// ---------------------------------------------------------------------------
struct AccI {
    typedef int32_t T;
    static T add(T a, T b) { return (T)((uint32_t)a + (uint32_t)b); }
    static T sub(T a, T b) { return (T)((uint32_t)a - (uint32_t)b); }
    static T mul(T a, int k) { return (T)((uint32_t)a * (uint32_t)k); }
    // C++ '/' truncates toward zero; divisors here are positive.
    static T div(T a, int d) { return a / d; }
    static T zero() { return 0; }
};

struct AccD {
    typedef double T;
    static T add(T a, T b) { return a + b; }
    static T sub(T a, T b) { return a - b; }
    static T mul(T a, int k) { return a * (double)k; }
    static T div(T a, int d) { return a / (double)d; }
    static T zero() { return 0.0; }
};

// Stored pyramid pixel <-> accumulator.  Storing an int into Int16/Int32 is a
// static_cast (wrap), VIGRA RequiresExplicitCast<integral>::cast(int).
template <class P> struct Store;
template <> struct Store<int16_t> {
    typedef AccI A;
    static int32_t load(int16_t v) { return v; }
    static int16_t wrap(int32_t v) { return (int16_t)(uint16_t)(uint32_t)v; }
    // NumericTraits<Int16>::fromPromote(int): clamp.
    static int16_t sat(int32_t v) { return (int16_t)std::min(32767, std::max(-32768, v)); }
};
template <> struct Store<int32_t> {
    typedef AccI A;
    static int32_t load(int32_t v) { return v; }
    static int32_t wrap(int32_t v) { return v; }
    static int32_t sat(int32_t v) { return v; }
};
template <> struct Store<double> {
    typedef AccD A;
    static double load(double v) { return v; }
    static double wrap(double v) { return v; }
    static double sat(double v) { return v; }
};

// ---------------------------------------------------------------------------
// VIGRA NumericTraits<T>::fromRealPromote (round half away from zero via
// v -/+ 0.5 then truncate, clamp).  NaN: x86-64 cvttsd2si yields INT_MIN for a
// 32-bit target; an Int16 target takes the low 16 bits of that, i.e. 0.
// ---------------------------------------------------------------------------
template <class T> T fromRealPromote(double v);
template <> inline int32_t fromRealPromote<int32_t>(double v) {
    if (isnan(v)) return INT32_MIN;
    if (v < 0.0) return v < (double)INT32_MIN ? INT32_MIN : (int32_t)(v - 0.5);
    return v > (double)INT32_MAX ? INT32_MAX : (int32_t)(v + 0.5);
}
template <> inline int16_t fromRealPromote<int16_t>(double v) {
    if (isnan(v)) return 0;
    if (v < 0.0) return v < -32768.0 ? (int16_t)-32768 : (int16_t)(int32_t)(v - 0.5);
    return v > 32767.0 ? (int16_t)32767 : (int16_t)(int32_t)(v + 0.5);
}
template <> inline uint8_t fromRealPromote<uint8_t>(double v) {
    if (isnan(v)) return 0;
    if (v < 0.0) return 0;
    return v > 255.0 ? (uint8_t)255 : (uint8_t)(int32_t)(v + 0.5);
}
template <> inline uint16_t fromRealPromote<uint16_t>(double v) {
    if (isnan(v)) return 0;
    if (v < 0.0) return 0;
    return v > 65535.0 ? (uint16_t)65535 : (uint16_t)(int32_t)(v + 0.5);
}
template <> inline double fromRealPromote<double>(double v) { return v; }
template <> inline float fromRealPromote<float>(double v) { return (float)v; }

// ---------------------------------------------------------------------------
// reduce, one channel.
// src/pyramid.h:166-573 (alpha) and :601-848 (no alpha).
// ---------------------------------------------------------------------------
template <class P>
void reduce_alpha_t(bool wrap, const P* src, int sw, int sh, long sp,
                    const uint8_t* al, long ap_,
                    P* dst, int dw, int dh, long dp, uint8_t* da, long dap)
{
    typedef Store<P> St;
    typedef typename St::A A;
    typedef typename A::T S;
    std::vector<S> h(dw), c0(dw), c1(dw), cp(dw);
    std::vector<int> m(dw), m0(dw), m1(dw), mp(dw);

    // Horizontal row filter with alpha-masked samples.
    auto row_filter = [&](int y) {
        const P* r = src + (long)y * sp;
        const uint8_t* a = al + (long)y * ap_;
        auto V = [&](int x) -> S { return a[x] ? (S)St::load(r[x]) : A::zero(); };
        auto M = [&](int x) -> int { return a[x] ? 1 : 0; };
        S r0 = A::zero(), r1 = A::zero(), rp = A::zero();
        int q0 = 0, q1 = 0, qp = 0;
        if (wrap) {
            r0 = V(sw - 2); rp = A::mul(V(sw - 1), 4);
            q0 = M(sw - 2); qp = 4 * M(sw - 1);
        }
        r1 = A::add(r0, rp); r0 = V(0);
        q1 = q0 + qp; q0 = M(0);
        int j = 0;
        for (int x = 1; x < sw; ++x) {
            if (x & 1) {
                rp = A::mul(V(x), 4); qp = 4 * M(x);
            } else {
                h[j] = A::add(A::add(A::add(r1, A::mul(r0, 6)), rp), V(x));
                m[j] = q1 + 6 * q0 + qp + M(x);
                r1 = A::add(r0, rp); r0 = V(x);
                q1 = q0 + qp; q0 = M(x);
                ++j;
            }
        }
        if (sw & 1) {
            S t = A::add(r1, A::mul(r0, 6));
            int qt = q1 + 6 * q0;
            if (wrap) {
                t = A::add(A::add(t, A::mul(V(0), 4)), V(1));
                qt += 4 * M(0) + M(1);
            }
            h[j] = t; m[j] = qt;
        } else {
            S t = A::add(A::add(r1, A::mul(r0, 6)), rp);
            int qt = q1 + 6 * q0 + qp;
            if (wrap) { t = A::add(t, V(0)); qt += M(0); }
            h[j] = t; m[j] = qt;
        }
    };

    auto emit = [&](int oy, int j, S ip, int apsum) {
        if (apsum) {
            dst[(long)oy * dp + j] = St::wrap(A::div(ip, apsum));
            da[(long)oy * dap + j] = 255;
        } else {
            dst[(long)oy * dp + j] = (P)0;
            da[(long)oy * dap + j] = 0;
        }
    };

    row_filter(0);
    for (int j = 0; j < dw; ++j) {
        c1[j] = A::zero(); c0[j] = h[j]; cp[j] = A::zero();
        m1[j] = 0; m0[j] = m[j]; mp[j] = 0;
    }
    for (int y = 1; y < sh; ++y) {
        row_filter(y);
        if (y & 1) {
            for (int j = 0; j < dw; ++j) { cp[j] = A::mul(h[j], 4); mp[j] = 4 * m[j]; }
        } else {
            for (int j = 0; j < dw; ++j) {
                S ip = A::add(A::add(c1[j], A::mul(c0[j], 6)), cp[j]);
                int apsum = m1[j] + 6 * m0[j] + mp[j];
                c1[j] = A::add(c0[j], cp[j]); c0[j] = h[j];
                m1[j] = m0[j] + mp[j]; m0[j] = m[j];
                ip = A::add(ip, c0[j]); apsum += m0[j];
                emit(y / 2 - 1, j, ip, apsum);
            }
        }
    }
    const int oy = (sh - 1) / 2;
    for (int j = 0; j < dw; ++j) {
        S ip; int apsum;
        if (sh & 1) {
            ip = A::add(c1[j], A::mul(c0[j], 6)); apsum = m1[j] + 6 * m0[j];
        } else {
            ip = A::add(A::add(c1[j], A::mul(c0[j], 6)), cp[j]); apsum = m1[j] + 6 * m0[j] + mp[j];
        }
        emit(oy, j, ip, apsum);
    }
    (void)dh;
}

template <class P>
void reduce_plain_t(bool wrap, const P* src, int sw, int sh, long sp,
                    P* dst, int dw, int dh, long dp)
{
    typedef Store<P> St;
    typedef typename St::A A;
    typedef typename A::T S;
    std::vector<S> h(dw), c0(dw), c1(dw), cp(dw);

    auto row_filter = [&](int y) {
        const P* r = src + (long)y * sp;
        auto X = [&](int x) -> S { return (S)St::load(r[x]); };
        S r0, r1, rp;
        if (wrap) { r0 = X(sw - 2); rp = A::mul(X(sw - 1), 4); }
        else { r0 = X(0); rp = A::mul(X(0), 4); }
        r1 = A::add(r0, rp); r0 = X(0);
        int j = 0;
        for (int x = 1; x < sw; ++x) {
            if (x & 1) {
                rp = A::mul(X(x), 4);
            } else {
                h[j] = A::add(A::add(A::add(r1, A::mul(r0, 6)), rp), X(x));
                r1 = A::add(r0, rp); r0 = X(x);
                ++j;
            }
        }
        if (sw & 1) {
            if (wrap) h[j] = A::add(A::add(A::add(r1, A::mul(r0, 6)), A::mul(X(0), 4)), X(1));
            else h[j] = A::add(r1, A::mul(r0, 11));
        } else {
            if (wrap) h[j] = A::add(A::add(A::add(r1, A::mul(r0, 6)), rp), X(0));
            else h[j] = A::add(A::add(A::add(r1, A::mul(r0, 6)), rp), A::div(rp, 4));
        }
    };

    row_filter(0);
    for (int j = 0; j < dw; ++j) { c0[j] = h[j]; c1[j] = A::mul(c0[j], 5); cp[j] = A::zero(); }
    for (int y = 1; y < sh; ++y) {
        row_filter(y);
        if (y & 1) {
            for (int j = 0; j < dw; ++j) cp[j] = A::mul(h[j], 4);
        } else {
            for (int j = 0; j < dw; ++j) {
                S ip = A::add(A::add(c1[j], A::mul(c0[j], 6)), cp[j]);
                c1[j] = A::add(c0[j], cp[j]); c0[j] = h[j];
                ip = A::add(ip, c0[j]);
                dst[(long)(y / 2 - 1) * dp + j] = St::wrap(A::div(ip, 256));
            }
        }
    }
    const int oy = (sh - 1) / 2;
    for (int j = 0; j < dw; ++j) {
        S ip;
        if (sh & 1) ip = A::add(c1[j], A::mul(c0[j], 11));
        else ip = A::add(A::add(A::add(c1[j], A::mul(c0[j], 6)), cp[j]), A::div(cp[j], 4));
        dst[(long)oy * dp + j] = St::wrap(A::div(ip, 256));
    }
    (void)dh;
}

// ---------------------------------------------------------------------------
// expand, one channel.  src/pyramid.h:873-1419.
// ---------------------------------------------------------------------------

// Horizontal weight sum behind the SKIPSM_* scale constants for destination
// column X (src/pyramid.h:1199-1411).  A 1-pixel-wide source always uses the
// no-wraparound constants.
inline int expand_hsum(int X, int sw, bool wrap, bool dst_even)
{
    const int j = X >> 1;
    if (sw == 1) return (X & 1) ? 4 : 6;
    if ((X & 1) == 0) {
        if (!wrap) return 6 + (j > 0 ? 1 : 0) + (j < sw - 1 ? 1 : 0);
        if (dst_even) return 8;
        return 6 + (j == 0 ? 4 : 1) + (j == sw - 1 ? 4 : 1);
    }
    if (!wrap) return 4 + (j < sw - 1 ? 4 : 0);
    return 8;
}

inline int expand_vsum(int Y, int sh)
{
    const int i = Y >> 1;
    if ((Y & 1) == 0) return 6 + (i > 0 ? 1 : 0) + (i < sh - 1 ? 1 : 0);
    return 1 + (i < sh - 1 ? 1 : 0);
}

template <class P>
void expand_t(bool add, bool wrap, const P* src, int sw, int sh, long sp,
              P* dst, int dw, int dh, long dp)
{
    typedef Store<P> St;
    typedef typename St::A A;
    typedef typename A::T S;
    const bool dst_even = (dw & 1) == 0;
    std::vector<S> hA(sw), hB(sw), c0a(sw), c0b(sw), c1a(sw), c1b(sw);

    auto row_filter = [&](int i) {
        const P* r = src + (long)i * sp;
        auto X = [&](int x) -> S { return (S)St::load(r[x]); };
        S r0 = X(0);
        S r1 = A::zero();
        if (wrap) { r1 = X(sw - 1); if (!dst_even) r1 = A::mul(r1, 4); }
        for (int s = 1; s < sw; ++s) {
            hA[s - 1] = A::add(A::add(r1, A::mul(r0, 6)), X(s));
            hB[s - 1] = A::mul(A::add(r0, X(s)), 4);
            r1 = r0; r0 = X(s);
        }
        const int j = sw - 1;
        if (wrap && (i == 0 || sw > 1)) {
            if (dst_even) {
                hA[j] = A::add(A::add(r1, A::mul(r0, 6)), X(0));
            } else {
                hA[j] = A::add(A::add(r1, A::mul(r0, 6)), A::mul(X(0), 4));
            }
            hB[j] = A::mul(A::add(r0, X(0)), 4);
        } else {
            hA[j] = A::add(r1, A::mul(r0, 6));
            hB[j] = A::mul(r0, 4);
        }
    };

    auto put = [&](int X, int Y, S o) {
        if (X >= dw || Y >= dh) return;
        const int d = expand_hsum(X, sw, wrap, dst_even) * expand_vsum(Y, sh);
        o = A::div(o, d);
        P& t = dst[(long)Y * dp + X];
        if (add) t = St::sat(A::add((S)St::load(t), o));
        else t = St::wrap(A::sub((S)St::load(t), o));
    };

    row_filter(0);
    for (int j = 0; j < sw; ++j) {
        c0a[j] = hA[j]; c0b[j] = hB[j]; c1a[j] = A::zero(); c1b[j] = A::zero();
    }
    for (int i = 1; i <= sh; ++i) {
        const bool last = (i == sh);
        if (!last) row_filter(i);
        const int Y = 2 * (i - 1);
        for (int j = 0; j < sw; ++j) {
            S o00, o10, o01, o11;
            if (!last) {
                o00 = A::add(A::add(c1a[j], A::mul(c0a[j], 6)), hA[j]);
                o10 = A::add(A::add(c1b[j], A::mul(c0b[j], 6)), hB[j]);
                o01 = A::add(c0a[j], hA[j]);
                o11 = A::add(c0b[j], hB[j]);
            } else {
                o00 = A::add(c1a[j], A::mul(c0a[j], 6));
                o10 = A::add(c1b[j], A::mul(c0b[j], 6));
                o01 = c0a[j];
                o11 = c0b[j];
            }
            put(2 * j, Y, o00);
            put(2 * j + 1, Y, o10);
            put(2 * j, Y + 1, o01);
            put(2 * j + 1, Y + 1, o11);
            if (!last) {
                c1a[j] = c0a[j]; c1b[j] = c0b[j]; c0a[j] = hA[j]; c0b[j] = hB[j];
            }
        }
    }
}

inline int check_reduce_sizes(int sw, int sh, int dw, int dh)
{
    if (sw < 2 || sh < 2) return -1;   // "src image too small in reduce"
    if (dw != (sw + 1) / 2 || dh != (sh + 1) / 2) return -2;
    return 0;
}

inline int check_expand_sizes(int sw, int sh, int dw, int dh)
{
    if (sw < 1 || sh < 1) return -1;
    if (dw != 2 * sw && dw != 2 * sw - 1) return -2;
    if (dh != 2 * sh && dh != 2 * sh - 1) return -2;
    if (dw < 2 && dh < 2) return -2;
    return 0;
}

// ---------------------------------------------------------------------------
// Converters.  src/fixmath.h:116-206 and :209-358.
// ---------------------------------------------------------------------------
// Deterministic dither draw shared with the GPU (the reference draws GSL
// MT19937 seeded from clock(), src/mersenne.cc:79-94): one 32-bit hash per
// pixel keyed by (seed, x, y), two rounds of the murmur3 32-bit finaliser;
// channel c takes bits [10c, 10c + frac - 2] as its offset on the
// reference's grid (r & half_m1) + quarter (src/fixmath.h:246-264).
inline uint32_t fmix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    return h ^ (h >> 16);
}

inline int dither_draw(uint64_t seed, int x, int y, int c, int frac_bits)
{
    const uint32_t s = (uint32_t)seed ^ (uint32_t)(seed >> 32);
    const uint32_t h = fmix32(s ^ fmix32(((uint32_t)y * 0x9E3779B1u) ^ (uint32_t)x));
    return (int)((h >> (10 * c)) & ((1u << (frac_bits - 1)) - 1u)) + (1 << (frac_bits - 2));
}



template <class I, class P>
void to_pyramid_int(int frac, int C, const I* img, int w, int h, long ip, P* out, long pp, long pl)
{
#pragma omp parallel for schedule(guided)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < C; ++c)
                out[c * pl + (long)y * pp + x] = (P)((P)img[(long)y * ip + (long)x * C + c] << frac);
}

inline void to_pyramid_log(int C, const float* img, int w, int h, long ip, double* out, long pp, long pl)
{
#pragma omp parallel for schedule(guided)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < C; ++c) {
                const float v = img[(long)y * ip + (long)x * C + c];
                out[c * pl + (long)y * pp + x] = v >= 0.0 ? 1.0 + log1p((double)v) : 1.0 / (1.0 - (double)v);
            }
}

template <class P>
void to_pyramid_fixed(int frac, int C, const float* img, int w, int h, long ip, P* out, long pp, long pl)
{
    const double scale = (double)(1u << frac);
#pragma omp parallel for schedule(guided)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < C; ++c)
                out[c * pl + (long)y * pp + x] = fromRealPromote<P>((double)img[(long)y * ip + (long)x * C + c] * scale);
}

template <class P, class I>
void from_pyramid_int(int frac, int C, const P* pl0, int w, int h, long pp, long pl,
                      const uint8_t* mask, long mp, I* img, long ip, int dither, uint64_t seed)
{
    const int32_t quarter = 1 << (frac - 2);
    const int32_t three_quarter = 3 << (frac - 2);
    const int32_t fmask = (1 << frac) - 1;
    const int32_t imax = (int32_t)std::numeric_limits<I>::max();
#pragma omp parallel for schedule(guided)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            if (!mask[(long)y * mp + x]) continue;
            for (int c = 0; c < C; ++c) {
                const int32_t v = pl0[c * pl + (long)y * pp + x];
                const int32_t q = v >> frac;
                const int32_t f = v & fmask;
                int32_t r;
                if (f >= quarter && f < three_quarter) {
                    if (dither == ORC_DITHER_DET) {
                        r = dither_draw(seed, x, y, c, frac) <= f ? q + 1 : q;
                    } else {
                        r = f >= (1 << (frac - 1)) ? q + 1 : q;   // round half up
                    }
                } else if (f >= quarter) {
                    r = q + 1;
                } else {
                    r = q;
                }
                img[(long)y * ip + (long)x * C + c] = (I)std::min(imax, std::max(0, r));
            }
        }
}

inline void from_pyramid_log(int C, const double* pl0, int w, int h, long pp, long pl,
                      const uint8_t* mask, long mp, float* img, long ip)
{
#pragma omp parallel for schedule(guided)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            if (!mask[(long)y * mp + x]) continue;
            for (int c = 0; c < C; ++c) {
                const double v = pl0[c * pl + (long)y * pp + x];
                img[(long)y * ip + (long)x * C + c] = (float)(v >= 1.0 ? expm1(v - 1.0) : 1.0 - 1.0 / v);
            }
        }
}

// CartesianBlendFunctor, src/blend.h:41-76.
template <class M, class P>
void blend_t(int C, int w, int h, const M* mask, long mp, const P* white, long wp, long wl,
             P* black, long bp, long bl, double white_value)
{
#pragma omp parallel for schedule(guided)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const double wc = (double)mask[(long)y * mp + x] / white_value;
            if (wc >= 1.0) {
                for (int c = 0; c < C; ++c) black[c * bl + (long)y * bp + x] = white[c * wl + (long)y * wp + x];
                continue;
            }
            if (wc <= 0.0) continue;
            const double bc = 1.0 - wc;
            for (int c = 0; c < C; ++c) {
                P& b = black[c * bl + (long)y * bp + x];
                const double v = (wc * (double)white[c * wl + (long)y * wp + x]) + (bc * (double)b);
                b = fromRealPromote<P>(v);
            }
        }
}

// ---------------------------------------------------------------------------
// Packed planar pyramid: level l holds C planes of w_l*h_l (pitch w_l).
// ---------------------------------------------------------------------------
struct Dims { int w, h; long off; };

inline std::vector<Dims> pyramid_dims(int w, int h, int levels, int C, long* total)
{
    std::vector<Dims> d;
    long off = 0;
    for (int l = 0; l < levels; ++l) {
        d.push_back({w, h, off});
        off += (long)C * w * h;
        w = (w + 1) >> 1;
        h = (h + 1) >> 1;
    }
    if (total) *total = off;
    return d;
}

template <class P>
int reduce_dispatch(bool wrap, const P* s, int sw, int sh, long sp, const uint8_t* a, long ap,
                    P* d, int dw, int dh, long dp, uint8_t* da, long dap)
{
    int rc = check_reduce_sizes(sw, sh, dw, dh);
    if (rc) return rc;
    if (a) reduce_alpha_t<P>(wrap, s, sw, sh, sp, a, ap, d, dw, dh, dp, da, dap);
    else reduce_plain_t<P>(wrap, s, sw, sh, sp, d, dw, dh, dp);
    return 0;
}

// gaussianPyramid with alpha (src/pyramid.h:1467-1559) or without
// (src/pyramid.h:1562-1636) into a packed planar buffer whose level 0 is
// already filled.
template <class P>
int gaussian_levels(P* buf, const std::vector<Dims>& d, int C, bool wrap, const uint8_t* alpha0)
{
    std::vector<uint8_t> a_prev, a_next;
    for (size_t l = 1; l < d.size(); ++l) {
        const Dims& s = d[l - 1];
        const Dims& t = d[l];
        const uint8_t* sa = nullptr;
        if (alpha0) {
            sa = (l == 1) ? alpha0 : a_prev.data();
            a_next.assign((size_t)t.w * t.h, 0);
        }
        for (int c = 0; c < C; ++c) {
            int rc = reduce_dispatch<P>(wrap, buf + s.off + (long)c * s.w * s.h, s.w, s.h, s.w,
                                        sa, s.w,
                                        buf + t.off + (long)c * t.w * t.h, t.w, t.h, t.w,
                                        alpha0 ? a_next.data() : nullptr, t.w);
            if (rc) return rc;
        }
        if (alpha0) a_prev.swap(a_next);
    }
    return 0;
}

// laplacianPyramid's expand-subtract loop (src/pyramid.h:1696-1708).
template <class P>
void laplacian_levels(P* buf, const std::vector<Dims>& d, int C, bool wrap)
{
    for (size_t l = 0; l + 1 < d.size(); ++l) {
        const Dims& f = d[l];
        const Dims& c = d[l + 1];
        for (int ch = 0; ch < C; ++ch)
            expand_t<P>(false, wrap, buf + c.off + (long)ch * c.w * c.h, c.w, c.h, c.w,
                        buf + f.off + (long)ch * f.w * f.h, f.w, f.h, f.w);
    }
}

// collapsePyramid (src/pyramid.h:1757-1784).
template <class P>
void collapse_levels(P* buf, const std::vector<Dims>& d, int C, bool wrap)
{
    for (int l = (int)d.size() - 2; l >= 0; --l) {
        const Dims& f = d[l];
        const Dims& c = d[l + 1];
        for (int ch = 0; ch < C; ++ch)
            expand_t<P>(true, wrap, buf + c.off + (long)ch * c.w * c.h, c.w, c.h, c.w,
                        buf + f.off + (long)ch * f.w * f.h, f.w, f.h, f.w);
    }
}

template <class I> constexpr int itype_code() { return sizeof(I) == 1 ? ORC_U8 : ORC_U16; }

// copyToPyramidImage (src/fixmath.h:2018-2115): an RGB image in the CIECAM02
// blend space enters as JCh (ConvertVectorToJCHPyramidFunctor), anything
// else through ConvertScalarToPyramidFunctor.
template <class I, class P>
void image_to_level0(int frac, int C, const I* img, int w, int h, P* out)
{
    if (C == 3 && orc_blend_colorspace() == ORC_COLORSPACE_CIECAM) {
        orc_rgb_to_jch_level0(itype_code<I>(), w, h, img, out, (long)w * h);
        return;
    }
    to_pyramid_int<I, P>(frac, C, img, w, h, (long)w * C, out, w, (long)w * h);
}

inline void image_to_level0(int, int C, const float* img, int w, int h, double* out)
{
    to_pyramid_log(C, img, w, h, (long)w * C, out, w, (long)w * h);
}

// copyFromPyramidImageIf (src/fixmath.h:2157-2294), JCh likewise.
template <class P, class I>
void level0_to_image(int frac, int C, const P* lv, int w, int h, const uint8_t* mask, I* img,
                     int dither, uint64_t seed)
{
    if (C == 3 && orc_blend_colorspace() == ORC_COLORSPACE_CIECAM) {
        orc_jch_level0_to_rgb_if(itype_code<I>(), w, h, lv, (long)w * h, mask, img);
        return;
    }
    from_pyramid_int<P, I>(frac, C, lv, w, h, w, (long)w * h, mask, w, img, (long)w * C, dither, seed);
}

inline void level0_to_image(int, int C, const double* lv, int w, int h, const uint8_t* mask, float* img,
                     int, uint64_t)
{
    from_pyramid_log(C, lv, w, h, w, (long)w * h, mask, w, img, (long)w * C);
}

// One enblend blend step (src/enblend.h:378-510) for image type I with
// pyramid type P and mask pyramid type M.
template <class I, class P, class M>
int enblend_pair_t(int C, int w, int h, int levels, bool wrap, int img_frac, int mask_frac,
                   const I* white, const uint8_t* walpha, I* black, uint8_t* balpha,
                   const uint8_t* mask, int dither, uint64_t seed)
{
    long nimg = 0, nmask = 0;
    std::vector<Dims> di = pyramid_dims(w, h, levels, C, &nimg);
    std::vector<Dims> dm = pyramid_dims(w, h, levels, 1, &nmask);

    // Gaussian pyramid of the seam mask, no alpha.
    std::vector<M> mgp((size_t)nmask);
    to_pyramid_int<uint8_t, M>(mask_frac, 1, mask, w, h, w, mgp.data(), w, (long)w * h);
    int rc = gaussian_levels<M>(mgp.data(), dm, 1, wrap, nullptr);
    if (rc) return rc;

    std::vector<P> wlp((size_t)nimg), blp((size_t)nimg);
    image_to_level0(img_frac, C, white, w, h, wlp.data());
    rc = gaussian_levels<P>(wlp.data(), di, C, wrap, walpha);
    if (rc) return rc;
    laplacian_levels<P>(wlp.data(), di, C, wrap);

    image_to_level0(img_frac, C, (const I*)black, w, h, blp.data());
    rc = gaussian_levels<P>(blp.data(), di, C, wrap, balpha);
    if (rc) return rc;
    laplacian_levels<P>(blp.data(), di, C, wrap);

    // Union alpha (src/enblend.h:461-463).
#pragma omp parallel for schedule(guided)
    for (long i = 0; i < (long)w * h; ++i)
        if (walpha[i]) balpha[i] = 255;

    // blend with white = 255 << mask_frac (src/enblend.h:472-474).
    const double white_value = (double)((M)(255 << mask_frac));
    for (int l = 0; l < levels; ++l)
        blend_t<M, P>(C, di[l].w, di[l].h, mgp.data() + dm[l].off, di[l].w,
                      wlp.data() + di[l].off, di[l].w, (long)di[l].w * di[l].h,
                      blp.data() + di[l].off, di[l].w, (long)di[l].w * di[l].h, white_value);

    collapse_levels<P>(blp.data(), di, C, wrap);

    // copyFromPyramidImageIf with the union alpha (src/enblend.h:506-510).
    level0_to_image(img_frac, C, blp.data(), w, h, balpha, black, dither, seed);
    return 0;
}

template <class I, class P>
int pyramid_t(bool laplacian, int C, int w, int h, int levels, bool wrap, int frac,
              const I* img, const uint8_t* alpha, P* out)
{
    std::vector<Dims> d = pyramid_dims(w, h, levels, C, nullptr);
    image_to_level0(frac, C, img, w, h, out);
    int rc = gaussian_levels<P>(out, d, C, wrap, alpha);
    if (rc) return rc;
    if (laplacian) laplacian_levels<P>(out, d, C, wrap);
    return 0;
}


}  // namespace orc
