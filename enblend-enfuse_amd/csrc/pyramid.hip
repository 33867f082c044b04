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
#include "pyr_common.h"
#include "reduce_stream.h"
#include "small_levels.h"

namespace enbl {
namespace {

// ---------------------------------------------------------------------------
// reduce (streaming strips: reduce_stream.h)
// ---------------------------------------------------------------------------
// AW: alpha rows 2-byte aligned (no wraparound only).
template <class P, int C, bool ALPHA, bool WRAP, bool AW>
__global__ __launch_bounds__(256) void reduce_kernel(ReduceArgs a)
{
    const int wx = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
    if constexpr (WRAP) reduce_strip<P, C, ALPHA, true, PlaneRows<P>>(a, wx, blockIdx.y * a.rows);
    else reduce_strip_ring<P, C, ALPHA, PlaneSrc<P, C>, AW>(a, wx, blockIdx.y * a.rows);
}

// Level 0 read from the interleaved integer image, (P)v << frac on the fly
// (ConvertScalarToPyramidFunctor, src/fixmath.h:116-163).  WIDE: image rows
// aligned for whole-word pair loads; AW: alpha rows 2-byte aligned.
template <class I, class P, int C, bool ALPHA, bool WIDE, bool AW>
__global__ __launch_bounds__(256) void reduce_image_kernel(ReduceArgs a)
{
    const int wx = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
    reduce_strip_ring<P, C, ALPHA, IntImageSrc<I, P, C, WIDE>, AW>(a, wx, blockIdx.y * a.rows);
}

// Level 0 read from the interleaved float image, 1 + log1p(v) or
// 1 / (1 - v) in double on the fly (src/fixmath.h:165-186): the reduce of
// a float image's level 0 without storing that level.
// W0: the converted level 0 is written to a.conv as well.
template <int C, bool ALPHA, bool AW, bool W0>
__global__ __launch_bounds__(256) void reduce_float_image_kernel(ReduceArgs a)
{
    const int wx = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
    reduce_strip_ring<double, C, ALPHA, FloatImageSrc<C>, AW, W0>(a, wx, blockIdx.y * a.rows);
}

// Level 0 -> 1 of a mask Gaussian pyramid under the union alpha, level 0
// formed on the fly from the weight and norm planes (NormMaskSrc).
template <class M, bool AW>
__global__ __launch_bounds__(256) void reduce_norm_mask_kernel(NormReduceArgs a)
{
    const int wx = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
    reduce_strip_ring<M, 1, true, NormMaskSrc<M>, AW>(a, wx, blockIdx.y * a.rows);
}

// ---------------------------------------------------------------------------
// expand
// ---------------------------------------------------------------------------

template <class S, int C, bool WRAP>
struct ExpandRow {
    S x[C];
    S xw[WRAP ? C : 1];   // far-side sample of a wraparound edge lane
};

template <class S, int C>
struct DestPair {
    S d0[C], d1[C], e0[C], e1[C];   // rows Y and Y+1, columns X and X+1
};

template <class P, int C, bool ADD, bool WRAP>
struct ExpandLane {
    typedef Pix<P> Px;
    typedef typename Px::A A;
    typedef typename Px::S S;
    typedef ExpandRow<S, C, WRAP> Row;
    typedef DestPair<S, C> Dest;

    const ExpandArgs& a;
    int j;
    bool valid, owned, dst_even, xb;

    __device__ __forceinline__ void load(int i, Row& R) const {
        const P* src = static_cast<const P*>(a.src);
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            const P* row = src + ch * a.splane + (long)i * a.spitch;
            R.x[ch] = valid ? Px::load(row[j]) : A::zero();
            if (WRAP) {
                R.xw[ch] = A::zero();
                if (valid && j == 0) R.xw[ch] = Px::load(row[a.sw - 1]);
                else if (valid && j == a.sw - 1) R.xw[ch] = Px::load(row[0]);
            }
        }
    }

    __device__ __forceinline__ void load_dest(int Y, Dest& D) const {
        const P* dst = static_cast<const P*>(a.dst);
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            D.d0[ch] = D.d1[ch] = D.e0[ch] = D.e1[ch] = A::zero();
            if (owned) {
                const P* r0 = dst + ch * a.dplane + (long)Y * a.dpitch + 2 * j;
                Vec2<P>::load(r0, xb, D.d0[ch], D.d1[ch]);
                if (Y + 1 < a.dh) Vec2<P>::load(r0 + a.dpitch, xb, D.e0[ch], D.e1[ch]);
            }
        }
    }

    // Polyphase row sums of source row i: hA for destination column 2j
    // (taps 1 6 1), hB for 2j+1 (taps 4 4); literal boundary forms of
    // src/pyramid.h:1148-1194 and the SKIPSM_EXPAND*_END macros.
    __device__ __forceinline__ void sums(const Row& R, int i, S (&hA)[C], S (&hB)[C]) const {
        const int sw = a.sw;
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            const S x = R.x[ch];
            const S l = lane_up(x);
            const S r = lane_down(x);
            if (sw == 1) {
                S r1 = A::zero();
                if (WRAP) r1 = dst_even ? x : A::mul(x, 4);
                if (WRAP && i == 0) {
                    hA[ch] = dst_even ? A::add(A::add(r1, A::mul(x, 6)), x) : A::add(A::add(r1, A::mul(x, 6)), A::mul(x, 4));
                    hB[ch] = A::mul(A::add(x, x), 4);
                } else {
                    hA[ch] = A::add(r1, A::mul(x, 6));
                    hB[ch] = A::mul(x, 4);
                }
            } else if (j == 0) {
                S r1 = A::zero();
                if (WRAP) r1 = dst_even ? R.xw[ch] : A::mul(R.xw[ch], 4);
                hA[ch] = A::add(A::add(r1, A::mul(x, 6)), r);
                hB[ch] = A::mul(A::add(x, r), 4);
            } else if (j == sw - 1) {
                if (WRAP) {
                    const S xf = R.xw[ch];
                    hA[ch] = dst_even ? A::add(A::add(l, A::mul(x, 6)), xf) : A::add(A::add(l, A::mul(x, 6)), A::mul(xf, 4));
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
    }
};

template <class P, int C, bool ADD, bool WRAP>
__global__ __launch_bounds__(256) void expand_kernel(ExpandArgs a)
{
    typedef ExpandLane<P, C, ADD, WRAP> Lane;
    typedef typename Lane::A A;
    typedef typename Lane::S S;
    typedef typename Lane::Row Row;
    typedef typename Lane::Dest Dest;
    typedef Pix<P> Px;

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
    const bool dst_even = (dw & 1) == 0;
    const int X = 2 * j;
    const bool xb = X + 1 < dw;
    const Lane L{a, j, valid, owned, dst_even, xb};
    P* dst = static_cast<P*>(a.dst);

    const int hdA = expand_hsum(X, sw, WRAP, dst_even);
    const int hdB = expand_hsum(X + 1, sw, WRAP, dst_even);

    S pA[C], pB[C], cA[C], cB[C], nA[C], nB[C];
#pragma unroll
    for (int ch = 0; ch < C; ++ch) { pA[ch] = pB[ch] = cA[ch] = cB[ch] = nA[ch] = nB[ch] = A::zero(); }
    Row rn, rnn;
    Dest dc, dn;
    {
        Row t;
        if (i0 > 0) {
            L.load(i0 - 1, t);
            L.sums(t, i0 - 1, pA, pB);
        }
        L.load(i0, t);
        L.sums(t, i0, cA, cB);
    }
    if (i0 + 1 < sh) L.load(i0 + 1, rn);
    L.load_dest(2 * i0, dc);

    for (int i = i0; i < i1; ++i) {
        const bool more = i + 1 < sh;
        if (i + 1 < i1) {
            if (i + 2 < sh) L.load(i + 2, rnn);
            L.load_dest(2 * (i + 1), dn);
        }
        if (more) L.sums(rn, i + 1, nA, nB);
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
                const S e0 = A::div(o00, hdA * v0);
                const S e1 = A::div(o10, hdB * v0);
                P w0, w1;
                if (ADD) { w0 = Px::sat(A::add(dc.d0[ch], e0)); w1 = Px::sat(A::add(dc.d1[ch], e1)); }
                else { w0 = Px::wrap(A::sub(dc.d0[ch], e0)); w1 = Px::wrap(A::sub(dc.d1[ch], e1)); }
                Vec2<P>::store(row0, xb, w0, w1);
                if (yb) {
                    const S f0 = A::div(o01, hdA * v1);
                    const S f1 = A::div(o11, hdB * v1);
                    if (ADD) { w0 = Px::sat(A::add(dc.e0[ch], f0)); w1 = Px::sat(A::add(dc.e1[ch], f1)); }
                    else { w0 = Px::wrap(A::sub(dc.e0[ch], f0)); w1 = Px::wrap(A::sub(dc.e1[ch], f1)); }
                    Vec2<P>::store(row0 + a.dpitch, xb, w0, w1);
                }
            }
            pA[ch] = cA[ch]; pB[ch] = cB[ch];
            cA[ch] = nA[ch]; cB[ch] = nB[ch];
        }
        rn = rnn;
        dc = dn;
    }
}

// ---------------------------------------------------------------------------
// Small levels: one thread per output, direct loads (small_levels.h).
//
// A coarse level has too few rows for the streaming kernels to hide the
// latency of their row-serial loop; here every thread fetches its whole
// window at once.
// ---------------------------------------------------------------------------
template <class P, int C, bool ALPHA, bool WRAP>
__global__ __launch_bounds__(256) void reduce_small_kernel(ReduceArgs a)
{
    reduce_small_body<P, C, ALPHA, WRAP>(a, blockIdx.x, blockIdx.y);
}

template <class P, int C, bool ADD, bool WRAP>
__global__ __launch_bounds__(256) void expand_small_kernel(ExpandArgs a)
{
    expand_small_body<P, C, ADD, WRAP>(a, blockIdx.x, blockIdx.y);
}

// Levels of at most this many (reduce: destination, expand: source) pixels
// take the direct kernels.
constexpr long kSmallLevel = 1L << 20;

// Rows of output per strip: enough strips to keep every SIMD busy, but long
// enough that the 3-row (reduce) / 2-row (expand) strip halo stays small.
inline int strip_rows(int waves_x, int out_rows)
{
    const long target = 8192;
    long r = ((long)waves_x * out_rows + target - 1) / target;
    if (r < 4) r = 4;
    if (r > 32) r = 32;
    return (int)r;
}

template <class P, int C, bool ALPHA>
void launch_reduce_t(ReduceArgs a, hipStream_t s)
{
    if (a.force_small == ENBL_KERNELS_DIRECT || (a.force_small == ENBL_KERNELS_AUTO && (long)a.dw * a.dh <= kSmallLevel)) {
        dim3 grid((a.dw + 63) / 64, (a.dh + 3) / 4);
        if (a.wrap) hipLaunchKernelGGL((reduce_small_kernel<P, C, ALPHA, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((reduce_small_kernel<P, C, ALPHA, false>), grid, dim3(256), 0, s, a);
        return;
    }
    const int waves_x = (a.dw + kOwned - 1) / kOwned + 1;
    if (a.rows <= 0) a.rows = strip_rows(waves_x, a.dh);
    dim3 block(256);
    dim3 grid((waves_x + 3) / 4, (a.dh + a.rows - 1) / a.rows);
    const bool aw = !ALPHA || (((uintptr_t)a.salpha & 1) == 0 && (a.sapitch & 1) == 0);
    if (a.wrap) hipLaunchKernelGGL((reduce_kernel<P, C, ALPHA, true, false>), grid, block, 0, s, a);
    else if (aw) hipLaunchKernelGGL((reduce_kernel<P, C, ALPHA, false, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((reduce_kernel<P, C, ALPHA, false, false>), grid, block, 0, s, a);
}

// Streaming reduce of an interleaved integer image (no wraparound).
template <class I, class P, int C, bool ALPHA>
void launch_reduce_image_t(ReduceArgs a, hipStream_t s)
{
    const int waves_x = (a.dw + kOwned - 1) / kOwned + 1;
    if (a.rows <= 0) a.rows = strip_rows(waves_x, a.dh);
    dim3 grid((waves_x + 3) / 4, (a.dh + a.rows - 1) / a.rows);
    const bool aw = !ALPHA || (((uintptr_t)a.salpha & 1) == 0 && (a.sapitch & 1) == 0);
    // pairs are whole words: 4-byte rows for 16-bit samples (and one 8-bit
    // channel: 2-byte rows); 8-bit RGB always loads per sample
    const uintptr_t am = sizeof(I) == 2 ? 3 : 1;
    const bool iw = (sizeof(I) == 2 || C == 1) && ((uintptr_t)a.src & am) == 0 && (a.spitch & 1) == 0;
    if (iw && aw) hipLaunchKernelGGL((reduce_image_kernel<I, P, C, ALPHA, true, true>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((reduce_image_kernel<I, P, C, ALPHA, false, false>), grid, dim3(256), 0, s, a);
}

template <int C, bool ALPHA>
void launch_reduce_float_image_t(ReduceArgs a, hipStream_t s)
{
    const int waves_x = (a.dw + kOwned - 1) / kOwned + 1;
    if (a.rows <= 0) a.rows = strip_rows(waves_x, a.dh);
    dim3 grid((waves_x + 3) / 4, (a.dh + a.rows - 1) / a.rows);
    const bool aw = !ALPHA || (((uintptr_t)a.salpha & 1) == 0 && (a.sapitch & 1) == 0);
    if (a.conv) {
        if (aw) hipLaunchKernelGGL((reduce_float_image_kernel<C, ALPHA, true, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((reduce_float_image_kernel<C, ALPHA, false, true>), grid, dim3(256), 0, s, a);
    } else {
        if (aw) hipLaunchKernelGGL((reduce_float_image_kernel<C, ALPHA, true, false>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((reduce_float_image_kernel<C, ALPHA, false, false>), grid, dim3(256), 0, s, a);
    }
}

template <class P, int C, bool ADD>
void launch_expand_t(ExpandArgs a, hipStream_t s)
{
    if (a.force_small == ENBL_KERNELS_DIRECT || (a.force_small == ENBL_KERNELS_AUTO && (long)a.sw * a.sh <= kSmallLevel)) {
        dim3 grid((a.sw + 63) / 64, (a.sh + 3) / 4);
        if (a.wrap) hipLaunchKernelGGL((expand_small_kernel<P, C, ADD, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((expand_small_kernel<P, C, ADD, false>), grid, dim3(256), 0, s, a);
        return;
    }
    const int waves_x = (a.sw + kOwned - 1) / kOwned + 1;
    if (a.rows <= 0) a.rows = strip_rows(waves_x, a.sh);
    dim3 block(256);
    dim3 grid((waves_x + 3) / 4, (a.sh + a.rows - 1) / a.rows);
    if (a.wrap) hipLaunchKernelGGL((expand_kernel<P, C, ADD, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((expand_kernel<P, C, ADD, false>), grid, block, 0, s, a);
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
    switch (ptype) {
    case ENBL_I16: return reduce_by_c<int16_t>(a, s);
    case ENBL_I32: return reduce_by_c<int32_t>(a, s);
    case ENBL_F64: return reduce_by_c<double>(a, s);
    }
    return ENBL_E_UNSUPPORTED;
}

int launch_reduce_image(int itype, int ptype, ReduceArgs a, hipStream_t s)
{
    if (a.wrap || (a.channels != 1 && a.channels != 3)) return ENBL_E_UNSUPPORTED;
    const bool alpha = a.salpha != nullptr;
#define ENBL_RI(I, P)                                                                                        \
    do {                                                                                                     \
        if (a.channels == 1) alpha ? launch_reduce_image_t<I, P, 1, true>(a, s) : launch_reduce_image_t<I, P, 1, false>(a, s); \
        else alpha ? launch_reduce_image_t<I, P, 3, true>(a, s) : launch_reduce_image_t<I, P, 3, false>(a, s); \
        return 0;                                                                                            \
    } while (0)
    if (itype == ENBL_U8 && ptype == ENBL_I16) ENBL_RI(uint8_t, int16_t);
    if (itype == ENBL_U8 && ptype == ENBL_I32) ENBL_RI(uint8_t, int32_t);
    if (itype == ENBL_U16 && ptype == ENBL_I32) ENBL_RI(uint16_t, int32_t);
#undef ENBL_RI
    if (itype == ENBL_F32 && ptype == ENBL_F64) {
        if (a.channels == 1) alpha ? launch_reduce_float_image_t<1, true>(a, s) : launch_reduce_float_image_t<1, false>(a, s);
        else alpha ? launch_reduce_float_image_t<3, true>(a, s) : launch_reduce_float_image_t<3, false>(a, s);
        return 0;
    }
    return ENBL_E_UNSUPPORTED;
}

int launch_reduce_norm_mask(int mtype, NormReduceArgs a, hipStream_t s)
{
    if (a.wrap || a.channels != 1 || !a.salpha) return ENBL_E_UNSUPPORTED;
    const int waves_x = (a.dw + kOwned - 1) / kOwned + 1;
    if (a.rows <= 0) a.rows = strip_rows(waves_x, a.dh);
    dim3 grid((waves_x + 3) / 4, (a.dh + a.rows - 1) / a.rows);
    const bool aw = ((uintptr_t)a.salpha & 1) == 0 && (a.sapitch & 1) == 0;
    if (mtype == ENBL_I32) {
        if (aw) hipLaunchKernelGGL((reduce_norm_mask_kernel<int32_t, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((reduce_norm_mask_kernel<int32_t, false>), grid, dim3(256), 0, s, a);
        return 0;
    }
    if (mtype == ENBL_I16) {
        if (aw) hipLaunchKernelGGL((reduce_norm_mask_kernel<int16_t, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((reduce_norm_mask_kernel<int16_t, false>), grid, dim3(256), 0, s, a);
        return 0;
    }
    return ENBL_E_UNSUPPORTED;
}

int launch_expand(int ptype, ExpandArgs a, hipStream_t s)
{
    switch (ptype) {
    case ENBL_I16: return expand_by_c<int16_t>(a, s);
    case ENBL_I32: return expand_by_c<int32_t>(a, s);
    case ENBL_F64: return expand_by_c<double>(a, s);
    }
    return ENBL_E_UNSUPPORTED;
}

}  // namespace enbl
