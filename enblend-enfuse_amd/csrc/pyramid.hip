// Burt-Adelson reduce and expand for gfx950.
//
// Reference: src/pyramid.h:166-573 (reduce with alpha), :601-848 (reduce
// without alpha), :873-1419 (expand).  The reference evaluates both with
// single-threaded SKIPSM raster passes.  Here each 64-lane wave owns a strip of
// 62 output columns and streams down the rows of its strip:
//   * every lane loads the source samples of its own column(s) with one
//     coalesced vector load per row and channel, and takes the horizontal
//     neighbours from the adjacent lanes through cross-lane shuffles;
//   * the horizontal 5-tap (reduce) or 3-tap polyphase (expand) row sums are
//     formed in registers and a rolling window of row sums gives the vertical
//     pass, so every source row is read from HBM once per strip;
//   * the boundary lanes evaluate the reference's literal boundary expressions
//     (clamp, zero-extension, wraparound and its 4x tap weights) from direct
//     loads.
// The operation order of every sum matches the reference's SKIPSM state
// updates, which makes the double path bit-identical to it and the integer
// path exact.
#include "enbl_device.h"
#include "enbl_internal.h"

namespace enbl {
namespace {

template <class P> struct Vec2;
template <> struct Vec2<int16_t> {
    __device__ static __forceinline__ void load(const int16_t* p, bool both, int32_t& a, int32_t& b) {
        if (both) {
            const uint32_t v = *reinterpret_cast<const uint32_t*>(p);
            a = (int16_t)(v & 0xffffu);
            b = (int16_t)(v >> 16);
        } else {
            a = p[0];
            b = 0;
        }
    }
    __device__ static __forceinline__ void store(int16_t* p, bool both, int16_t a, int16_t b) {
        if (both) {
            *reinterpret_cast<uint32_t*>(p) = (uint32_t)(uint16_t)a | ((uint32_t)(uint16_t)b << 16);
        } else {
            p[0] = a;
        }
    }
};
template <> struct Vec2<int32_t> {
    __device__ static __forceinline__ void load(const int32_t* p, bool both, int32_t& a, int32_t& b) {
        if (both) {
            const int2 v = *reinterpret_cast<const int2*>(p);
            a = v.x;
            b = v.y;
        } else {
            a = p[0];
            b = 0;
        }
    }
    __device__ static __forceinline__ void store(int32_t* p, bool both, int32_t a, int32_t b) {
        if (both) {
            *reinterpret_cast<int2*>(p) = make_int2(a, b);
        } else {
            p[0] = a;
        }
    }
};
template <> struct Vec2<double> {
    __device__ static __forceinline__ void load(const double* p, bool both, double& a, double& b) {
        if (both) {
            const double2 v = *reinterpret_cast<const double2*>(p);
            a = v.x;
            b = v.y;
        } else {
            a = p[0];
            b = 0.0;
        }
    }
    __device__ static __forceinline__ void store(double* p, bool both, double a, double b) {
        if (both) {
            *reinterpret_cast<double2*>(p) = make_double2(a, b);
        } else {
            p[0] = a;
        }
    }
};

__device__ __forceinline__ int wrap_index(int k, int n) {
    k %= n;
    return k < 0 ? k + n : k;
}

// ---------------------------------------------------------------------------
// reduce
// ---------------------------------------------------------------------------
template <class P, int C, bool ALPHA>
__global__ __launch_bounds__(256) void reduce_kernel(ReduceArgs a)
{
    typedef Pix<P> Px;
    typedef typename Px::A A;
    typedef typename Px::S S;

    const int lane = threadIdx.x & (kWave - 1);
    const int wx = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
    if (wx * kOwned - 1 >= a.dw) return;  // whole wave beyond the level
    const int i0 = blockIdx.y * a.rows;
    if (i0 >= a.dh) return;
    const int i1 = min(a.dh, i0 + a.rows);

    const int sw = a.sw, sh = a.sh;
    const int j = wx * kOwned + lane - 1;
    const int c = 2 * j;
    const bool valid = j >= 0 && j < a.dw;
    const bool owned = lane >= 1 && lane <= kOwned && valid;
    const bool both = c + 1 < sw;
    const bool edge = valid && (j == 0 || c + 2 >= sw);
    const bool wrap = a.wrap != 0;

    const P* src = static_cast<const P*>(a.src);
    P* dst = static_cast<P*>(a.dst);

    // Horizontal row sum of source row r at centre column c, all channels;
    // alpha weight sum in *hm.  Executed by every lane of the wave (shuffles).
    auto hrow = [&](int r, S (&h)[C], int& hm) {
        const uint8_t* arow = ALPHA ? a.salpha + (long)r * a.sapitch : nullptr;
        int m0 = 0, m1 = 0;
        if (ALPHA && valid) {
            m0 = arow[c] ? 1 : 0;
            m1 = both ? (arow[c + 1] ? 1 : 0) : 0;
        }
        int ml0 = 0, ml1 = 0, mr0 = 0;
        if (ALPHA) {
            ml0 = lane_up(m0);
            ml1 = lane_up(m1);
            mr0 = lane_down(m0);
        }
        if (ALPHA) {
            if (!edge) {
                hm = ml0 + 4 * ml1 + 6 * m0 + 4 * m1 + mr0;
            } else {
                auto M = [&](int k) -> int { return arow[k] ? 1 : 0; };
                int ma, mb, md, me;
                if (wrap) {
                    ma = M(wrap_index(c - 2, sw)); mb = M(wrap_index(c - 1, sw));
                    md = M(wrap_index(c + 1, sw)); me = M(wrap_index(c + 2, sw));
                    hm = ma + 4 * mb + 6 * m0 + 4 * md + me;
                } else {
                    ma = c >= 2 ? M(c - 2) : 0; mb = c >= 1 ? M(c - 1) : 0;
                    md = c + 1 < sw ? M(c + 1) : 0; me = c + 2 < sw ? M(c + 2) : 0;
                    hm = ma + 4 * mb + 6 * m0 + 4 * md + me;
                }
            }
        }
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            const P* row = src + ch * a.splane + (long)r * a.spitch;
            S x0 = A::zero(), x1 = A::zero();
            if (valid) {
                Vec2<P>::load(row + c, both, x0, x1);
                if (ALPHA) {
                    if (!m0) x0 = A::zero();
                    if (!m1) x1 = A::zero();
                }
            }
            const S l0 = lane_up(x0);
            const S l1 = lane_up(x1);
            const S r0 = lane_down(x0);
            if (!edge) {
                h[ch] = A::add(A::add(A::add(A::add(l0, A::mul(l1, 4)), A::mul(x0, 6)), A::mul(x1, 4)), r0);
                continue;
            }
            // Boundary lanes: literal boundary expressions of the reference.
            auto F = [&](int k) -> S {
                if (ALPHA) return arow[k] ? Px::load(row[k]) : A::zero();
                return Px::load(row[k]);
            };
            if (wrap) {
                const S va = F(wrap_index(c - 2, sw)), vb = F(wrap_index(c - 1, sw));
                const S vd = F(wrap_index(c + 1, sw)), ve = F(wrap_index(c + 2, sw));
                h[ch] = A::add(A::add(A::add(A::add(va, A::mul(vb, 4)), A::mul(x0, 6)), A::mul(vd, 4)), ve);
            } else if (ALPHA) {
                const S va = c >= 2 ? F(c - 2) : A::zero();
                const S vb = c >= 1 ? F(c - 1) : A::zero();
                const S a0 = A::add(va, A::mul(vb, 4));
                if (c == sw - 1) {
                    h[ch] = A::add(a0, A::mul(x0, 6));
                } else if (c == sw - 2) {
                    h[ch] = A::add(A::add(a0, A::mul(x0, 6)), A::mul(x1, 4));
                } else {
                    h[ch] = A::add(A::add(A::add(a0, A::mul(x0, 6)), A::mul(x1, 4)), F(c + 2));
                }
            } else {
                const S va = c >= 2 ? F(c - 2) : F(0);
                const S vb = c >= 1 ? F(c - 1) : F(0);
                const S a0 = A::add(va, A::mul(vb, 4));
                if (c == sw - 1) {
                    h[ch] = A::add(a0, A::mul(x0, 11));
                } else if (c == sw - 2) {
                    const S d4 = A::mul(x1, 4);
                    h[ch] = A::add(A::add(A::add(a0, A::mul(x0, 6)), d4), A::div(d4, 4));
                } else {
                    h[ch] = A::add(A::add(A::add(a0, A::mul(x0, 6)), A::mul(x1, 4)), F(c + 2));
                }
            }
        }
    };

    S hm2[C], hm1[C], hc[C], hp1[C], hp2[C];
    int qm2 = 0, qm1 = 0, qc = 0, qp1 = 0, qp2 = 0;
#pragma unroll
    for (int ch = 0; ch < C; ++ch) { hm2[ch] = A::zero(); hm1[ch] = A::zero(); }
    if (i0 > 0) {
        hrow(2 * i0 - 2, hm2, qm2);
        hrow(2 * i0 - 1, hm1, qm1);
    }
    hrow(2 * i0, hc, qc);

    for (int i = i0; i < i1; ++i) {
        const int r = 2 * i;
        if (r + 1 < sh) hrow(r + 1, hp1, qp1);
        if (r + 2 < sh) hrow(r + 2, hp2, qp2);

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
                const S a0 = r == 0 ? A::zero() : A::add(hm2[ch], A::mul(hm1[ch], 4));
                if (r == sh - 1) v = A::add(a0, A::mul(hc[ch], 6));
                else if (r == sh - 2) v = A::add(A::add(a0, A::mul(hc[ch], 6)), A::mul(hp1[ch], 4));
                else v = A::add(A::add(A::add(a0, A::mul(hc[ch], 6)), A::mul(hp1[ch], 4)), hp2[ch]);
            } else {
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
                dst[ch * a.dplane + (long)i * a.dpitch + j] = out;
            }
            hm2[ch] = hc[ch];
            hm1[ch] = hp1[ch];
            hc[ch] = hp2[ch];
        }
        if (ALPHA && owned) a.dalpha[(long)i * a.dapitch + j] = ap ? 255 : 0;
        qm2 = qc; qm1 = qp1; qc = qp2;
    }
}

// ---------------------------------------------------------------------------
// expand
// ---------------------------------------------------------------------------

// Horizontal weight sum behind the reference's SKIPSM_* scale constants for
// destination column X (src/pyramid.h:1199-1411).  A 1-pixel-wide source
// always takes the no-wraparound constants.
__device__ __forceinline__ int expand_hsum(int X, int sw, bool wrap, bool dst_even) {
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

__device__ __forceinline__ int expand_vsum(int Y, int sh) {
    const int i = Y >> 1;
    if ((Y & 1) == 0) return 6 + (i > 0 ? 1 : 0) + (i < sh - 1 ? 1 : 0);
    return 1 + (i < sh - 1 ? 1 : 0);
}

template <class P, int C, bool ADD>
__global__ __launch_bounds__(256) void expand_kernel(ExpandArgs a)
{
    typedef Pix<P> Px;
    typedef typename Px::A A;
    typedef typename Px::S S;

    const int lane = threadIdx.x & (kWave - 1);
    const int wx = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
    if (wx * kOwned - 1 >= a.sw) return;
    const int i0 = blockIdx.y * a.rows;
    if (i0 >= a.sh) return;
    const int i1 = min(a.sh, i0 + a.rows);

    const int sw = a.sw, sh = a.sh, dw = a.dw, dh = a.dh;
    const int j = wx * kOwned + lane - 1;
    const bool valid = j >= 0 && j < sw;
    const bool owned = lane >= 1 && lane <= kOwned && valid;
    const bool wrap = a.wrap != 0;
    const bool dst_even = (dw & 1) == 0;

    const P* src = static_cast<const P*>(a.src);
    P* dst = static_cast<P*>(a.dst);

    // Horizontal polyphase sums of source row i: hA for destination column 2j
    // (taps 1 6 1), hB for 2j+1 (taps 4 4).
    auto hrow = [&](int i, S (&hA)[C], S (&hB)[C]) {
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            const P* row = src + ch * a.splane + (long)i * a.spitch;
            const S x = valid ? Px::load(row[j]) : A::zero();
            const S l = lane_up(x);
            const S r = lane_down(x);
            if (sw == 1) {
                S r1 = A::zero();
                if (wrap) r1 = dst_even ? x : A::mul(x, 4);
                if (wrap && i == 0) {
                    hA[ch] = dst_even ? A::add(A::add(r1, A::mul(x, 6)), x)
                                      : A::add(A::add(r1, A::mul(x, 6)), A::mul(x, 4));
                    hB[ch] = A::mul(A::add(x, x), 4);
                } else {
                    hA[ch] = A::add(r1, A::mul(x, 6));
                    hB[ch] = A::mul(x, 4);
                }
            } else if (j == 0) {
                S r1 = A::zero();
                if (wrap) {
                    const S xl = Px::load(row[sw - 1]);
                    r1 = dst_even ? xl : A::mul(xl, 4);
                }
                hA[ch] = A::add(A::add(r1, A::mul(x, 6)), r);
                hB[ch] = A::mul(A::add(x, r), 4);
            } else if (j == sw - 1) {
                if (wrap) {
                    const S xf = Px::load(row[0]);
                    hA[ch] = dst_even ? A::add(A::add(l, A::mul(x, 6)), xf)
                                      : A::add(A::add(l, A::mul(x, 6)), A::mul(xf, 4));
                    hB[ch] = A::mul(A::add(x, xf), 4);
                } else {
                    hA[ch] = A::add(l, A::mul(x, 6));
                    hB[ch] = A::mul(x, 4);
                }
            } else {
                hA[ch] = A::add(A::add(l, A::mul(x, 6)), r);
                hB[ch] = A::mul(A::add(x, r), 4);
            }
        }
    };

    const int X = 2 * j;
    const bool xb = X + 1 < dw;
    const int hdA = expand_hsum(X, sw, wrap, dst_even);
    const int hdB = expand_hsum(X + 1, sw, wrap, dst_even);

    S pA[C], pB[C], cA[C], cB[C], nA[C], nB[C];
#pragma unroll
    for (int ch = 0; ch < C; ++ch) { pA[ch] = A::zero(); pB[ch] = A::zero(); }
    if (i0 > 0) hrow(i0 - 1, pA, pB);
    hrow(i0, cA, cB);

    for (int i = i0; i < i1; ++i) {
        const bool more = i + 1 < sh;
        if (more) hrow(i + 1, nA, nB);
        const int Y = 2 * i;
        const bool yb = Y + 1 < dh;
        const int v0 = expand_vsum(Y, sh);
        const int v1 = expand_vsum(Y + 1, sh);
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
            if (owned) {
                P* row0 = dst + ch * a.dplane + (long)Y * a.dpitch + X;
                S d0, d1;
                Vec2<P>::load(row0, xb, d0, d1);
                const S e0 = A::div(o00, hdA * v0);
                const S e1 = A::div(o10, hdB * v0);
                P w0, w1;
                if (ADD) { w0 = Px::sat(A::add(d0, e0)); w1 = Px::sat(A::add(d1, e1)); }
                else { w0 = Px::wrap(A::sub(d0, e0)); w1 = Px::wrap(A::sub(d1, e1)); }
                Vec2<P>::store(row0, xb, w0, w1);
                if (yb) {
                    P* row1 = row0 + a.dpitch;
                    Vec2<P>::load(row1, xb, d0, d1);
                    const S f0 = A::div(o01, hdA * v1);
                    const S f1 = A::div(o11, hdB * v1);
                    if (ADD) { w0 = Px::sat(A::add(d0, f0)); w1 = Px::sat(A::add(d1, f1)); }
                    else { w0 = Px::wrap(A::sub(d0, f0)); w1 = Px::wrap(A::sub(d1, f1)); }
                    Vec2<P>::store(row1, xb, w0, w1);
                }
            }
            pA[ch] = cA[ch]; pB[ch] = cB[ch];
            cA[ch] = nA[ch]; cB[ch] = nB[ch];
        }
    }
}

template <class P, int C, bool ALPHA>
void launch_reduce_t(const ReduceArgs& a, hipStream_t s)
{
    const int waves_x = (a.dw + kOwned - 1) / kOwned + 1;
    dim3 block(256);
    dim3 grid((waves_x + 3) / 4, (a.dh + a.rows - 1) / a.rows);
    hipLaunchKernelGGL((reduce_kernel<P, C, ALPHA>), grid, block, 0, s, a);
}

template <class P, int C, bool ADD>
void launch_expand_t(const ExpandArgs& a, hipStream_t s)
{
    const int waves_x = (a.sw + kOwned - 1) / kOwned + 1;
    dim3 block(256);
    dim3 grid((waves_x + 3) / 4, (a.sh + a.rows - 1) / a.rows);
    hipLaunchKernelGGL((expand_kernel<P, C, ADD>), grid, block, 0, s, a);
}

template <class P>
int reduce_by_c(const ReduceArgs& a, hipStream_t s)
{
    const bool alpha = a.salpha != nullptr;
    if (a.channels == 1) { alpha ? launch_reduce_t<P, 1, true>(a, s) : launch_reduce_t<P, 1, false>(a, s); return 0; }
    if (a.channels == 3) { alpha ? launch_reduce_t<P, 3, true>(a, s) : launch_reduce_t<P, 3, false>(a, s); return 0; }
    return ENBL_E_UNSUPPORTED;
}

template <class P>
int expand_by_c(const ExpandArgs& a, hipStream_t s)
{
    if (a.channels == 1) { a.add ? launch_expand_t<P, 1, true>(a, s) : launch_expand_t<P, 1, false>(a, s); return 0; }
    if (a.channels == 3) { a.add ? launch_expand_t<P, 3, true>(a, s) : launch_expand_t<P, 3, false>(a, s); return 0; }
    return ENBL_E_UNSUPPORTED;
}

}  // namespace

int launch_reduce(int ptype, ReduceArgs a, hipStream_t s)
{
    if (a.rows <= 0) a.rows = 16;
    switch (ptype) {
    case ENBL_I16: return reduce_by_c<int16_t>(a, s);
    case ENBL_I32: return reduce_by_c<int32_t>(a, s);
    case ENBL_F64: return reduce_by_c<double>(a, s);
    }
    return ENBL_E_UNSUPPORTED;
}

int launch_expand(int ptype, ExpandArgs a, hipStream_t s)
{
    if (a.rows <= 0) a.rows = 16;
    switch (ptype) {
    case ENBL_I16: return expand_by_c<int16_t>(a, s);
    case ENBL_I32: return expand_by_c<int32_t>(a, s);
    case ENBL_F64: return expand_by_c<double>(a, s);
    }
    return ENBL_E_UNSUPPORTED;
}

}  // namespace enbl
