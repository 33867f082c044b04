// Nearest-feature-transform seam mask (src/nearest.h:424-552, open
// boundaries) with the distance transform of the reference's OpenMP build:
// the Felzenszwalb-Huttenlocher 1-D lower-envelope transform in float over
// every column, then over every row of the intermediate image
// (src/openmp_vigra.h:244-470).
//
// A 1-D transform builds its lower envelope sequentially with a stack, and
// its float roundings (q^2 beyond 2^24, the envelope intersections, the
// final sum and the square root) decide which samples are nearest, so each
// column (first pass) and each row (second pass) is one thread running the
// reference's operations in the reference's order.  Both distance images
// (to the white-only and to the black-only pixels) run in the same launches.
// The stacks live in device scratch, entry k of thread t at k * T + t, so
// lanes at similar depths share cache lines; the top entry stays in
// registers.
#include <float.h>
#include <limits.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <type_traits>

#include "enbl_device.h"
#include "enbl_internal.h"

namespace enbl {
namespace {

struct NftArgs {
    const uint8_t* white; long wpitch;
    const uint8_t* black; long bpitch;
    int w, h, norm;
    float* inter; long iplane;     // 2 planes of w x h (pitch w): after the column pass
    float* dist; long dplane; long dpitch;   // 2 planes of distances (row pitch dpitch)
    int* sv; float* sz; float* sf;  // column stacks: sample index, envelope bound, sample value
    uint64_t* env; int* top;        // row envelopes, (w + 1) entries per row (env_entry); top index per row
    unsigned* tally;               // 2 counters of exclusive pixels
    uint8_t* mask; long mpitch;
};

// EuclideanTransform1D (src/openmp_vigra.h:304-357) of n samples: f(q) reads
// sample q, out(q, d) writes result q.  Stack entry k of this thread at
// index k * T (pointers already offset to the thread).
template <class F, class O>
__device__ __forceinline__ void euclid_1d(int n, F f, O out, int* v, float* z, float* fv, long T)
{
    const float inf = INFINITY;
    int k = 0;
    int vk = 0;                 // v[k], fv[k], z[k] of the top entry
    float fk = f(0), zk = -inf;
    v[0] = 0;
    fv[0] = fk;
    z[0] = zk;
    for (int q = 1; q < n; ++q) {
        const float fq = f(q);
        const float sum_q = fq + (float)q * (float)q;
        float s = (sum_q - (fk + (float)(vk * vk))) / (float)(2 * (q - vk));
        while (s <= zk) {
            --k;
            vk = v[k * T];
            fk = fv[k * T];
            zk = z[k * T];
            s = (sum_q - (fk + (float)(vk * vk))) / (float)(2 * (q - vk));
        }
        ++k;
        vk = q;
        fk = fq;
        zk = s;
        v[k * T] = q;
        fv[k * T] = fq;
        z[k * T] = s;
    }
    z[(long)(k + 1) * T] = inf;
    k = 0;
    vk = v[0];
    fk = fv[0];
    float zn = z[T];            // z[k + 1]
    for (int q = 0; q < n; ++q) {
        while (zn < (float)q) {
            ++k;
            vk = v[k * T];
            fk = fv[k * T];
            zn = z[(long)(k + 1) * T];
        }
        out(q, (float)((q - vk) * (q - vk)) + fk);
    }
}

// ManhattanTransform1D (src/openmp_vigra.h:283-301); `in(q)` reads the
// result written so far for the backward sweep.
template <class F, class O, class I>
__device__ __forceinline__ void manhattan_1d(int n, F f, O out, I in)
{
    float d = f(0);
    out(0, d);
    for (int q = 1; q < n; ++q) {
        d = fminf(f(q), d + 1.0f);
        out(q, d);
    }
    for (int q = n - 2; q >= 0; --q) {
        d = fminf(in(q), d + 1.0f);
        out(q, d);
    }
}

// Columns: t = blockIdx.y (0: white-only features, 1: black-only).
// f = FLT_MAX on background, 0 on a feature (src/openmp_vigra.h:384-392).
__global__ __launch_bounds__(64) void nft_columns_kernel(NftArgs a)
{
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int t = blockIdx.y;
    const bool live = x < a.w;
    const int xc = live ? x : a.w - 1;
    const uint8_t* p = t == 0 ? a.white : a.black;
    const uint8_t* q = t == 0 ? a.black : a.white;   // null: features are the non-zero samples of p
    const long pp = t == 0 ? a.wpitch : a.bpitch, qp = t == 0 ? a.bpitch : a.wpitch;
    unsigned count = 0;
    // saturating difference p - q (src/nearest.h:359-372)
    auto f = [&](int y) {
        const uint8_t u = *gmem(p + (long)y * pp + xc), v = q ? *gmem(q + (long)y * qp + xc) : (uint8_t)0;
        const bool feat = v < u;
        count += feat ? 1u : 0u;
        return feat ? 0.0f : FLT_MAX;
    };
    float* col = a.inter + t * a.iplane + xc;
    auto out = [&](int y, float d) { if (live) *gmem(col + (long)y * a.w) = d; };
    if (a.norm == 1) {
        manhattan_1d(a.h, f, out, [&](int y) { return *gmem(col + (long)y * a.w); });
    } else {
        const long T = (long)gridDim.y * gridDim.x * 64;
        const long tid = (long)t * gridDim.x * 64 + x;
        euclid_1d(a.h, f, out, a.sv + tid, a.sz + tid, a.sf + tid, T);
    }
    // each sample of the column was read once by f (forward sweep)
    count = live ? count : 0u;
    for (int o = 32; o > 0; o >>= 1) count += __shfl_xor(count, o);
    if (threadIdx.x == 0) atomicAdd(&a.tally[t], count);
}

// Euclidean column pass for columns of at most 4096 samples.  The column
// input is binary (0 on features, FLT_MAX elsewhere), and below 2^24 every
// float the 1-D transform forms from feature samples is exact: q^2, the
// intersections (q^2 - v^2) / (2 (q - v)) = (q + v) / 2, and the result
// (q - v)^2; background samples only ever give FLT_MAX + x = FLT_MAX and
// intersections of magnitude ~FLT_MAX, which never displace a feature.  So
// the transform equals the squared distance to the nearest feature of the
// column, or FLT_MAX in a column without features, which is found without a
// stack: the nearest feature at or above and the one at or below.
constexpr int kScanMax = 4096, kSeg = 64;

// Columns are cut into segments of kSeg rows, one lane per (CW columns,
// segment).  The first kernel finds each segment's first and last feature
// row (and counts the features for the overlap tally); the second turns
// them, per column, into the last feature above each segment and the first
// below it; the third carries those into the segment and runs both sweeps
// over its rows.  With CW = 4 a lane reads its four columns as one 32-bit
// word per row (rows 4-byte aligned); the columns past the last multiple of
// four take a CW = 1 launch of their own.
struct SegArgs {
    int* first;       // [transform][segment][column]: first feature row, or INT_MAX; then the first below
    int* last;        // last feature row, or -1; then the last above
    int nseg;
    int x0, x1;       // columns of this launch
};

// Feature bits of rows y0 .. y0 + kSeg - 1 of columns x .. x + CW - 1 (rows
// past h read as background).
template <int CW>
__device__ __forceinline__ void segment_bits(const NftArgs& a, int t, int x, int y0, uint64_t (&b)[CW])
{
    const uint8_t* p = t == 0 ? a.white : a.black;
    const uint8_t* q = t == 0 ? a.black : a.white;   // null: features are the non-zero samples of p
    const long pp = t == 0 ? a.wpitch : a.bpitch, qp = t == 0 ? a.bpitch : a.wpitch;
    using W = typename std::conditional<CW == 4, uint32_t, uint8_t>::type;
    W u[kSeg], v[kSeg];
#pragma unroll
    for (int i = 0; i < kSeg; ++i) {
        const bool in = y0 + i < a.h;
        u[i] = in ? *gmem((const W*)(p + (long)(y0 + i) * pp + x)) : (W)0;
        v[i] = in && q ? *gmem((const W*)(q + (long)(y0 + i) * qp + x)) : (W)0;
    }
#pragma unroll
    for (int c = 0; c < CW; ++c) {
        uint64_t m = 0;
#pragma unroll
        for (int i = 0; i < kSeg; ++i) {
            const unsigned uc = ((unsigned)u[i] >> (8 * c)) & 255u, vc = ((unsigned)v[i] >> (8 * c)) & 255u;
            m |= vc < uc ? (1ull << i) : 0ull;
        }
        b[c] = m;
    }
}

template <int CW>
__global__ __launch_bounds__(64) void nft_colseg_summary_kernel(NftArgs a, SegArgs g)
{
    const int x = g.x0 + (blockIdx.x * 64 + threadIdx.x) * CW;
    const int sg = blockIdx.y, t = blockIdx.z;
    const bool live = x < g.x1;
    const int xc = live ? x : g.x1 - CW;
    uint64_t b[CW];
    segment_bits<CW>(a, t, xc, sg * kSeg, b);
    unsigned count = 0;
    if (live) {
        const long o = ((long)t * g.nseg + sg) * a.w + x;
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            g.first[o + c] = b[c] ? sg * kSeg + __builtin_ctzll(b[c]) : INT_MAX;
            g.last[o + c] = b[c] ? sg * kSeg + 63 - __builtin_clzll(b[c]) : -1;
            count += (unsigned)__builtin_popcountll(b[c]);
        }
    }
    for (int o = 32; o > 0; o >>= 1) count += __shfl_xor(count, o);
    if (threadIdx.x == 0) atomicAdd(&a.tally[t], count);
}

// Per column: last[s] becomes the last feature row of the segments above s
// (-kNone if none) and first[s] the first feature row of the segments below
// s (h + kNone if none).
constexpr int kNone = 1 << 20, kSegMax = kScanMax / kSeg;

__global__ __launch_bounds__(64) void nft_colseg_carry_kernel(NftArgs a, SegArgs g)
{
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int t = blockIdx.y;
    if (x >= a.w) return;
    const long base = (long)t * g.nseg * a.w + x;
    int f[kSegMax], l[kSegMax];
#pragma unroll
    for (int s = 0; s < kSegMax; ++s)
        if (s < g.nseg) { f[s] = g.first[base + (long)s * a.w]; l[s] = g.last[base + (long)s * a.w]; }
    int above = -kNone;
#pragma unroll
    for (int s = 0; s < kSegMax; ++s)
        if (s < g.nseg) {
            const int ls = l[s];
            g.last[base + (long)s * a.w] = above;
            above = ls >= 0 ? ls : above;
        }
    int below = a.h + kNone;
#pragma unroll
    for (int s = kSegMax - 1; s >= 0; --s)
        if (s < g.nseg) {
            const int fs = f[s];
            g.first[base + (long)s * a.w] = below;
            below = fs != INT_MAX ? fs : below;
        }
}

// Row by row, the distance of each of the lane's columns to the last
// feature at or above (the segment's bits up to the row, else the carried
// one) and to the first at or below; with CW = 4 and rows of a multiple of
// four floats the four results leave as one 16-byte store.
template <int CW>
__global__ __launch_bounds__(64) void nft_colseg_apply_kernel(NftArgs a, SegArgs g)
{
    typedef float f4 __attribute__((ext_vector_type(4)));
    const int x = g.x0 + (blockIdx.x * 64 + threadIdx.x) * CW;
    const int sg = blockIdx.y, t = blockIdx.z;
    if (x >= g.x1) return;
    const int y0 = sg * kSeg;
    uint64_t b[CW];
    segment_bits<CW>(a, t, x, y0, b);
    const long o = ((long)t * g.nseg + sg) * a.w + x;
    int above[CW], below[CW];
#pragma unroll
    for (int c = 0; c < CW; ++c) { above[c] = g.last[o + c]; below[c] = g.first[o + c]; }
    float* col = a.inter + t * a.iplane + x;
    const bool wide = CW == 4 && (a.w & 3) == 0;
    const int rows = min(kSeg, a.h - y0);
    for (int i = 0; i < rows; ++i) {
        const int y = y0 + i;
        float d[CW];
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            const uint64_t up = b[c] & ((2ull << i) - 1ull);   // rows y0 .. y (all 64 at i = 63)
            const uint64_t dn = b[c] >> i;                      // rows y .. y0 + 63
            const int du = up ? i - (63 - __builtin_clzll(up)) : y - above[c];
            const int dd = dn ? __builtin_ctzll(dn) : below[c] - y;
            const int m = min(du, dd);
            d[c] = m >= kNone ? FLT_MAX : (float)(m * m);
        }
        float* r = col + (long)y * a.w;
        if constexpr (CW == 4) {
            if (wide) {
                *gmem((f4*)r) = f4{d[0], d[1], d[2], d[3]};
                continue;
            }
        }
#pragma unroll
        for (int c = 0; c < CW; ++c) *gmem(r + c) = d[c];
    }
}

// Manhattan rows of the intermediate image.
__global__ __launch_bounds__(64) void nft_rows_manhattan_kernel(NftArgs a)
{
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int t = blockIdx.y;
    if (y >= a.h) return;
    const float* row = a.inter + t * a.iplane + (long)y * a.w;
    float* drow = a.dist + t * a.dplane + (long)y * a.dpitch;
    manhattan_1d(a.w, [&](int x) { return *gmem(row + x); }, [&](int x, float d) { *gmem(drow + x) = d; },
                 [&](int x) { return *gmem(drow + x); });
}

// Euclidean row pass in two kernels.
//
// The lower envelope of a row is built by one lane running the reference's
// 1-D transform (src/openmp_vigra.h:304-357) in the reference's order; a wave
// holds only kEnvRows rows, so that its lanes rarely wait on each other's pops
// and the 2 h rows of the two transforms spread over many waves.  The stack's
// newest kEnvRing entries of every lane live in an LDS ring (entry e in slot
// e mod kEnvRing); every push is also written to the row's envelope in device
// memory as (sample index, envelope bound), and a pop below the ring reloads
// the kRefill entries under it in one batch, the sample values from the row
// itself.  The row samples come through an LDS tile of kEnvRows rows x kRT
// columns loaded with coalesced reads, the next tile already in flight; all
// 64 lanes prepare the tile, each one column: the sum f(q) + q^2 and the
// intersection with the parabola of q - 1, which is the stack top whenever q
// arrives, so a sample that pops nothing costs the serial lane no division.
// The entry under the top is kept in registers, so a sample that pops one
// entry reads nothing from LDS either.  Two runs of a tile take one step, each
// only after checking the condition that makes the reference's loop do
// exactly that: a run of FLT_MAX samples (columns without features) where each
// sample pops its predecessor and lands on the same entry, and a suffix whose
// intersections increase, where every sample pushes.  The envelope's final top
// index goes to a table.
//
// The output does not depend on the order in which the reference's forward
// output sweep visits the samples: the envelope bounds are finite beyond entry
// 0 (-inf) and strictly increasing (an entry is pushed only when its bound
// exceeds the one under it; every sample is finite, FLT_MAX at most), so the
// sweep's stop for sample q, the smallest k with z[k + 1] >= q, is the largest
// k with z[k] < q.  The second kernel finds it by bisection for the first of a
// run of kOutCols samples and walks up from there; every result
// sqrt((q - v_k)^2 + f(v_k)) is the same float.
constexpr int kEnvRows = 2, kEnvRing = 256, kRefill = 32, kRT = 64, kOutCols = 8;
// A tile's pushes (up to kRT) stay in the ring until they are written out.
static_assert(kEnvRing >= kRT && (kEnvRing & (kEnvRing - 1)) == 0, "ring size");

// Envelope entry: sample index in the low word, the bound's float bits in
// the high word.
__device__ __forceinline__ uint64_t env_entry(int v, float z)
{
    return ((uint64_t)(uint32_t)__float_as_int(z) << 32) | (uint32_t)v;
}
__device__ __forceinline__ int env_v(uint64_t e) { return (int)(uint32_t)e; }
__device__ __forceinline__ float env_z(uint64_t e) { return __int_as_float((int)(uint32_t)(e >> 32)); }

// (sum_q - s_v) / d for an even d = 2 (q - v) in [2, 2^24), correctly
// rounded like the IEEE division it replaces, in fewer dependent steps:
// y = RN(1/d) by one FMA Newton step from v_rcp_f32 (equal to 1.0f / d for
// every integer d below 2^24 on the device), q0 = RN(n y) within one ulp of
// n / d, r = n - d q0 exact by FMA, and RN(q0 + r y) = RN(n / d) (Markstein's
// correction, which holds whenever y = RN(1/d) and nothing under- or
// overflows; the sign of a zero quotient from n = -0 may differ, which
// sum_q >= 1 rules out here).  tests/test_gpu_nft.py checks both steps on
// the device through enbl_nft_quotients: y for every integer d below 2^24
// (every float significand), and n / d for every even d up to 2^16 and
// powers of two to 2^30 over numerators of every magnitude and sign.
__device__ __forceinline__ float env_div(float n, int d)
{
    const float fd = (float)d;
    const float y0 = __builtin_amdgcn_rcpf(fd);
    const float y = __builtin_fmaf(__builtin_fmaf(-fd, y0, 1.0f), y0, y0);
    const float q0 = n * y;
    return __builtin_fmaf(__builtin_fmaf(-q0, fd, n), y, q0);
}

__global__ __launch_bounds__(64) void nft_rows_env_kernel(NftArgs a)
{
    constexpr int R = kEnvRows, M = kEnvRing - 1, TP = kRT + 4;
    __shared__ uint4 ring[kEnvRing][R];   // (v, sum bits, z bits, 0)
    // Per row and tile column j: the sum f(q) + q^2 and the intersection with
    // the parabola of q - 1.
    __shared__ float tq[R][TP], t0[R][TP];
    __shared__ int span[R][2];
    __shared__ int bulk[R][2];   // a row's run of pushes: first tile column, stack index under it
    const int lane = threadIdx.x;
    const int t = blockIdx.y;
    const int row0 = blockIdx.x * R;
    const int nrows = min(R, a.h - row0);
    const int n = a.w;
    const bool own = lane < nrows;
    const int ro = own ? lane : 0;
    const float* __restrict__ src = a.inter + t * a.iplane + (long)row0 * n;
    const float* __restrict__ mine = src + (long)ro * n;
    uint64_t* __restrict__ env = a.env + ((long)t * a.h + row0 + ro) * (long)(n + 1);
    const float inf = INFINITY;
    float pf[R], carry[R];
    auto fetch = [&](int x0) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            pf[r] = r < nrows && x0 + lane < n ? *gmem(src + (long)r * n + x0 + lane) : 0.f;
    };
    // Top entry: index k, sample vk, sum_v = f(v) + v^2 (the reference's
    // fk + (float)(vk * vk), the same rounding of the same integer square),
    // bound zk.  The entry under it (k - 1) is kept in registers too while hb
    // is set (bv, bs, bz), so that a sample popping one entry reads nothing
    // from LDS.  Ring slots from lo up to k are valid.
    int k = 0, lo = 0;
    int vk = 0, bv = 0;
    float sk = 0.f, zk = -inf, bs = 0.f, bz = -inf;
    bool hb = false;
    int mk = -1;   // lowest stack index reached in this tile
    auto refill = [&]() {   // entries k - kRefill + 1 .. k into the ring, 8 at a time
        const int e0 = max(0, k - kRefill + 1);
#pragma unroll 1
        for (int c = k; c >= e0; c -= 8) {
            int tv[8];
            float tz[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int e = c - j;
                if (e >= e0) { const uint64_t p = *gmem(env + e); tv[j] = env_v(p); tz[j] = env_z(p); }
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int e = c - j;
                if (e >= e0) {
                    const float sv = *gmem(mine + tv[j]) + (float)tv[j] * (float)tv[j];
                    ring[e & M][lane] = make_uint4((unsigned)tv[j], __float_as_uint(sv), __float_as_uint(tz[j]), 0u);
                }
            }
        }
        lo = e0;
    };
#pragma unroll
    for (int r = 0; r < R; ++r) carry[r] = 0.f;
    fetch(0);
    for (int x0 = 0; x0 < n; x0 += kRT) {
        const int cn = min(kRT, n - x0);
        // Every lane prepares one column of the tile for all rows.
        const int c = x0 + lane;
#pragma unroll
        for (int r = 0; r < R; ++r) tq[r][lane] = pf[r] + (float)c * (float)c;
        __syncthreads();
        // Per row, from ballots over the columns: lb, the last column j >= 1
        // whose intersection does not exceed its left neighbour's (0 if none),
        // and fs, the first column from which every sample to the tile end is
        // FLT_MAX (a column without features in the column pass).
        int my_lb = 0, my_fs = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const float prev = lane ? tq[r][lane - 1] : carry[r];
            const float tz = (tq[r][lane] - prev) * 0.5f;   // RN(x / 2) either way
            t0[r][lane] = tz;
            carry[r] = tq[r][kRT - 1];
            const float tzp = __shfl_up(tz, 1);
            const unsigned long long brk = __ballot(lane >= 1 && lane < cn && !(tz > tzp));
            const unsigned long long fin = __ballot(lane < cn && tq[r][lane] != FLT_MAX);
            const int lbv = brk ? 63 - __builtin_clzll(brk) : 0;
            const int fsv = fin ? 64 - __builtin_clzll(fin) : 0;
            my_lb = lane == r ? lbv : my_lb;
            my_fs = lane == r ? fsv : my_fs;
        }
        __syncthreads();
        if (x0 + kRT < n) fetch(x0 + kRT);
        if (own) {
            int i = 0;
            if (x0 == 0) {   // sample 0 starts the envelope
                vk = 0; sk = tq[lane][0]; zk = -inf; hb = false;
                ring[0][lane] = make_uint4(0u, __float_as_uint(sk), __float_as_uint(-inf), 0u);
                i = 1;
            }
            int pst = -1, pk = 0;
            const int my_ff = min(my_fs, my_lb);
            float sum_n = tq[lane][i], s0_n = t0[lane][i];
            for (; i < cn; ++i) {
                const int q = x0 + i;
                const float sum_q = sum_n;
                const float s = s0_n;   // against q - 1, the top whenever q arrives
                sum_n = tq[lane][i + 1];   // TP > kRT: the read past the last column is never used
                s0_n = t0[lane][i + 1];
                if (i >= my_ff) {
                    if (i >= my_fs && sk == FLT_MAX && hb && zk >= 0.f) {
                        // Samples q .. b are FLT_MAX, and so is the top q - 1:
                        // each one meets the top at (FLT_MAX - FLT_MAX) / 2 = 0
                        // <= zk, pops it and meets the entry under it, A, at
                        // RN((FLT_MAX - sum_A) / (2 (p - v_A))) >= 0, which
                        // decreases with p, so if it still exceeds z_A at b
                        // every sample pushes onto A: A, b with the bound at b
                        // remain.
                        const int b = x0 + cn - 1;
                        const float sb = env_div(FLT_MAX - bs, 2 * (b - bv));
                        if (sb > bz) {
                            mk = min(mk, k - 1);
                            vk = b; sk = FLT_MAX; zk = sb;
                            ring[k & M][lane] = make_uint4((unsigned)b, __float_as_uint(FLT_MAX), __float_as_uint(sb), 0u);
                            break;
                        }
                    }
                    if (i >= my_lb && s > zk) {
                        // q pops nothing, and every later intersection of the
                        // tile exceeds its predecessor's: q .. b all push (their
                        // ring slots are written below by the whole wave).
                        pst = i;
                        pk = k;
                        const int last = cn - 1;
                        if (last > i) {
                            bv = x0 + last - 1; bs = tq[lane][last - 1]; bz = t0[lane][last - 1];
                        } else {
                            bv = vk; bs = sk; bz = zk;
                        }
                        hb = true;
                        vk = x0 + last; sk = tq[lane][last]; zk = t0[lane][last];
                        k += cn - i;
                        lo = max(lo, k - kEnvRing + 1);
                        break;
                    }
                }
                // The common cases without branches: q pops nothing (p0), or
                // pops the top and lands on the entry under it (p1, known in
                // registers).  Anything else runs the reference's loop.
                const float s1 = env_div(sum_q - bs, 2 * (q - bv));
                bool p0 = s > zk;
                const bool p1 = !p0 && hb && s1 > bz;
                float zq = p0 ? s : s1;
                if (!p0 && !p1) {
                    --k;
                    if (hb) {
                        vk = bv; sk = bs; zk = bz;
                        hb = false;
                    } else {
                        if (k < lo) refill();
                        const uint4 e = ring[k & M][lane];
                        vk = (int)e.x;
                        sk = __uint_as_float(e.y);
                        zk = __uint_as_float(e.z);
                    }
                    float ss = env_div(sum_q - sk, 2 * (q - vk));
                    while (ss <= zk) {   // never below entry 0, whose bound is -inf
                        --k;
                        if (k < lo) refill();
                        const uint4 e = ring[k & M][lane];
                        vk = (int)e.x;
                        sk = __uint_as_float(e.y);
                        zk = __uint_as_float(e.z);
                        ss = env_div(sum_q - sk, 2 * (q - vk));
                    }
                    mk = min(mk, k);
                    zq = ss;
                    p0 = true;   // q goes onto the entry reached
                }
                // p0: q goes onto the top, which becomes the entry under it;
                // p1: q replaces the top, over the same entry under it.
                mk = p0 ? mk : min(mk, k - 1);
                bv = p0 ? vk : bv;
                bs = p0 ? sk : bs;
                bz = p0 ? zk : bz;
                hb = true;
                k += p0 ? 1 : 0;
                lo = k - lo >= kEnvRing ? k - kEnvRing + 1 : lo;   // entry k - kEnvRing leaves the ring
                ring[k & M][lane] = make_uint4((unsigned)q, __float_as_uint(sum_q), __float_as_uint(zq), 0u);
                vk = q; sk = sum_q; zk = zq;
            }
            bulk[lane][0] = pst;
            bulk[lane][1] = pk;
            span[lane][0] = mk;
            span[lane][1] = k;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j0 = bulk[r][0];
            if (r < nrows && j0 >= 0 && lane >= j0 && lane < cn)
                ring[(bulk[r][1] + 1 + lane - j0) & M][r] = make_uint4((unsigned)(x0 + lane), __float_as_uint(tq[r][lane]),
                                                                    __float_as_uint(t0[r][lane]), 0u);
        }
        __syncthreads();
        // The entries this tile pushed, (mk, k] (at most kRT, all still in the
        // ring), go to the row's envelope in device memory, one coalesced
        // store per row; a refill only ever reads entries of earlier tiles.
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int e = span[r][0] + 1 + lane;
            if (r < nrows && e <= span[r][1]) {
                const uint4 v = ring[e & M][r];
                *gmem(a.env + ((long)t * a.h + row0 + r) * (long)(n + 1) + e) = env_entry((int)v.x, __uint_as_float(v.z));
            }
        }
        mk = k;
        __syncthreads();
    }
    if (own) *gmem(a.top + (long)t * a.h + row0 + lane) = k;
}

// Results of kOutCols consecutive samples per lane of one row (blockIdx.x),
// 256 x kOutCols samples per block, leaving through LDS (a lane's run padded
// by one word against bank conflicts) as coalesced stores.
__global__ __launch_bounds__(256) void nft_rows_out_kernel(NftArgs a)
{
    __shared__ float res[256 * (kOutCols + 1)];
    const int y = blockIdx.x, t = blockIdx.z;
    const int n = a.w;
    const int c0 = blockIdx.y * 256 * kOutCols;
    const int q0 = c0 + threadIdx.x * kOutCols;
    const long row = (long)t * a.h + y;
    const uint64_t* __restrict__ env = a.env + row * (long)(n + 1);
    const float* __restrict__ f = a.inter + t * a.iplane + (long)y * n;
    if (q0 < n) {
        const int K = *gmem(a.top + row);
        const float fq0 = (float)q0;
        int lo = 0, hi = K;   // largest k in [0, K] with z[k] < q0 (z[0] = -inf)
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (env_z(*gmem(env + mid)) < fq0) lo = mid;
            else hi = mid - 1;
        }
        int k = lo;
        int vk = env_v(*gmem(env + k));
        float fk = *gmem(f + vk);
        float zn = k < K ? env_z(*gmem(env + k + 1)) : INFINITY;
#pragma unroll
        for (int j = 0; j < kOutCols; ++j) {
            const int q = q0 + j;
            if (q < n) {
                while (zn < (float)q) {
                    ++k;
                    vk = env_v(*gmem(env + k));
                    fk = *gmem(f + vk);
                    zn = k < K ? env_z(*gmem(env + k + 1)) : INFINITY;
                }
                res[threadIdx.x * (kOutCols + 1) + j] = sqrtf((float)((q - vk) * (q - vk)) + fk);
            }
        }
    }
    __syncthreads();
    float* __restrict__ drow = a.dist + t * a.dplane + (long)y * a.dpitch;
#pragma unroll
    for (int j = 0; j < kOutCols; ++j) {
        const int c = j * 256 + threadIdx.x;
        if (c0 + c < n) *gmem(drow + c0 + c) = res[c / kOutCols * (kOutCols + 1) + c % kOutCols];
    }
}

// mask = dist12 < dist21 ? 255 : 0 (src/nearest.h:540-545)
__global__ __launch_bounds__(256) void nft_mask_kernel(NftArgs a)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= a.w) return;
    const long i = (long)y * a.dpitch + x;
    *gmem(a.mask + (long)y * a.mpitch + x) = *gmem(a.dist + i) < *gmem(a.dist + a.dplane + i) ? 255 : 0;
}

// quadruple_image (src/nearest.h:138-185): the w x h alpha tiled tx x ty
// times into a W x H image (pitch W).
__global__ __launch_bounds__(256) void nft_tile_kernel(const uint8_t* __restrict__ src, long sp, int w, int h,
                                                       uint8_t* __restrict__ dst, int W)
{
    const int X = blockIdx.x * 256 + threadIdx.x, Y = blockIdx.y;
    if (X >= W) return;
    const int x = X >= w ? X - w : X, y = Y >= h ? Y - h : Y;
    *gmem(dst + (long)Y * W + X) = *gmem(src + (long)y * sp + x);
}

// quater_image (src/nearest.h:200-285) fused with the mask: a doubled axis
// of length n takes position p + n for p < n / 2, p otherwise.
__global__ __launch_bounds__(256) void nft_mask_periodic_kernel(NftArgs a, int w, int h, int hx, int hy)
{
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const int X = hx && x < w / 2 ? x + w : x, Y = hy && y < h / 2 ? y + h : y;
    const long i = (long)Y * a.dpitch + X;
    *gmem(a.mask + (long)y * a.mpitch + x) = *gmem(a.dist + i) < *gmem(a.dist + a.dplane + i) ? 255 : 0;
}

__global__ void nft_zero_kernel(unsigned* p) { if (threadIdx.x < 2) p[threadIdx.x] = 0u; }

// The envelope quotient of the row pass over arrays, so that its equality
// with the IEEE division can be checked on the device it runs on.
__global__ __launch_bounds__(256) void nft_quotient_kernel(const float* __restrict__ num, const int* __restrict__ den,
                                                           float* __restrict__ out, long count)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < count) *gmem(out + i) = env_div(*gmem(num + i), *gmem(den + i));
}

}  // namespace
}  // namespace enbl

using namespace enbl;
using namespace enbl::host;

namespace {

// The column and row passes of nt (1 or 2) distance transforms.
int nft_passes(enbl_context* ctx, NftArgs& a, int nt)
{
    const int w = a.w, h = a.h;
    const int gx_c = (w + 63) / 64, gx_r = (h + 63) / 64;
    const size_t plane = (size_t)w * h;
    // Euclidean stacks, one buffer for both passes (the column stacks are dead
    // when the row pass starts): n + 1 entries of 12 bytes per column thread
    // (only for columns longer than kScanMax), n + 1 entries of 8 bytes per
    // row plus the row's top index
    const size_t entries = (size_t)gx_c * 64 * nt * ((size_t)h + 1);
    const size_t col_bytes = h > kScanMax ? entries * 12 : 0;
    const size_t env_bytes = (size_t)nt * h * ((size_t)w + 1) * sizeof(uint64_t);
    const size_t row_bytes = env_bytes + (size_t)nt * h * sizeof(int);
    void *fbuf, *sbuf = nullptr, *tbuf;
    int rc = scratch(ctx, 16, (size_t)nt * plane * sizeof(float), &fbuf);
    if (!rc && a.norm == 2) rc = scratch(ctx, 17, std::max(col_bytes, row_bytes), &sbuf);
    if (!rc) rc = scratch(ctx, 18, 64, &tbuf);
    if (rc) return rc;
    a.inter = (float*)fbuf; a.iplane = (long)plane;
    a.sv = nullptr; a.sz = nullptr; a.sf = nullptr; a.env = nullptr; a.top = nullptr;
    if (a.norm == 2) {
        a.sv = (int*)sbuf;
        a.sz = (float*)((char*)sbuf + entries * 4);
        a.sf = (float*)((char*)sbuf + entries * 8);
        a.env = (uint64_t*)sbuf;
        a.top = (int*)((char*)sbuf + env_bytes);
    }
    a.tally = (unsigned*)tbuf;
    {
        Timed t(ctx, "nft_columns", nt * (double)plane * (2.0 + 4.0));
        hipLaunchKernelGGL(nft_zero_kernel, dim3(1), dim3(64), 0, ctx->stream, a.tally);
        if (a.norm == 2 && h <= kScanMax) {
            SegArgs g;
            g.nseg = (h + kSeg - 1) / kSeg;
            void* sb;
            rc = scratch(ctx, 20, 2 * (size_t)nt * g.nseg * w * sizeof(int), &sb);
            if (rc) return rc;
            g.first = (int*)sb;
            g.last = (int*)sb + (size_t)nt * g.nseg * w;
            // word reads need every row of both alphas 4-byte aligned
            const auto al4 = [](const void* p, long pitch) { return p == nullptr || (((uintptr_t)p & 3) == 0 && (pitch & 3) == 0); };
            const int w4 = al4(a.white, a.wpitch) && al4(a.black, a.bpitch) ? w & ~3 : 0;
            SegArgs g4 = g, g1 = g;
            g4.x0 = 0; g4.x1 = w4;
            g1.x0 = w4; g1.x1 = w;
            const unsigned gx4 = (unsigned)((w4 / 4 + 63) / 64), gx1 = (unsigned)((w - w4 + 63) / 64);
            if (gx4) hipLaunchKernelGGL(nft_colseg_summary_kernel<4>, dim3(gx4, g.nseg, nt), dim3(64), 0, ctx->stream, a, g4);
            if (gx1) hipLaunchKernelGGL(nft_colseg_summary_kernel<1>, dim3(gx1, g.nseg, nt), dim3(64), 0, ctx->stream, a, g1);
            hipLaunchKernelGGL(nft_colseg_carry_kernel, dim3(gx_c, nt), dim3(64), 0, ctx->stream, a, g);
            if (gx4) hipLaunchKernelGGL(nft_colseg_apply_kernel<4>, dim3(gx4, g.nseg, nt), dim3(64), 0, ctx->stream, a, g4);
            if (gx1) hipLaunchKernelGGL(nft_colseg_apply_kernel<1>, dim3(gx1, g.nseg, nt), dim3(64), 0, ctx->stream, a, g1);
        } else {
            hipLaunchKernelGGL(nft_columns_kernel, dim3(gx_c, nt), dim3(64), 0, ctx->stream, a);
        }
        rc = launch_status("nft_columns");
        if (rc) return rc;
    }
    if (a.norm != 2) {
        Timed t(ctx, "nft_rows", nt * (double)plane * 8.0);
        hipLaunchKernelGGL(nft_rows_manhattan_kernel, dim3(gx_r, nt), dim3(64), 0, ctx->stream, a);
        return launch_status("nft_rows");
    }
    {
        // reads the row samples, writes at most w + 1 envelope entries per row
        Timed t(ctx, "nft_rows_env", nt * (double)plane * 12.0);
        hipLaunchKernelGGL(nft_rows_env_kernel, dim3((h + kEnvRows - 1) / kEnvRows, nt), dim3(64), 0, ctx->stream, a);
        rc = launch_status("nft_rows_env");
        if (rc) return rc;
    }
    {
        // reads the envelope and the samples it names, writes the distances
        Timed t(ctx, "nft_rows_out", nt * (double)plane * 16.0);
        const int gy = (w + 256 * kOutCols - 1) / (256 * kOutCols);
        hipLaunchKernelGGL(nft_rows_out_kernel, dim3(h, gy, nt), dim3(256), 0, ctx->stream, a);
        rc = launch_status("nft_rows_out");
    }
    return rc;
}

}  // namespace

extern "C" {

int enbl_distance_transform(enbl_context* ctx, int w, int h, const uint8_t* src, long spitch, int norm, float* dst,
                            long dpitch)
{
    if (!ctx || !src || !dst) return fail(ENBL_E_ARG, "enbl_distance_transform: null argument");
    if (w < 1 || h < 1 || spitch < w || dpitch < w) return fail(ENBL_E_SIZE, "enbl_distance_transform: bad size or pitch");
    if (norm != 1 && norm != 2) return fail(ENBL_E_UNSUPPORTED, "enbl_distance_transform: norm %d (1 Manhattan, 2 Euclidean)", norm);
    NftArgs a;
    memset(&a, 0, sizeof a);
    a.white = src; a.wpitch = spitch; a.black = nullptr; a.bpitch = 0;
    a.w = w; a.h = h; a.norm = norm;
    a.dist = dst; a.dplane = 0; a.dpitch = dpitch;
    return nft_passes(ctx, a, 1);
}

int enbl_nft_quotients(enbl_context* ctx, const float* num, const int* den, float* out, long count)
{
    if (!ctx || (count > 0 && (!num || !den || !out))) return fail(ENBL_E_ARG, "enbl_nft_quotients: null argument");
    if (count < 0) return fail(ENBL_E_SIZE, "enbl_nft_quotients: negative count");
    if (count == 0) return ENBL_OK;
    const long blocks = (count + 255) / 256;
    if (blocks > 0x7fffffffL) return fail(ENBL_E_SIZE, "enbl_nft_quotients: count too large");
    hipLaunchKernelGGL(nft_quotient_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, num, den, out, count);
    return launch_status("nft_quotients");
}

int enbl_nft_mask(enbl_context* ctx, int w, int h, const uint8_t* white, long wpitch, const uint8_t* black, long bpitch,
                  int norm, uint8_t* mask, long mpitch)
{
    if (!ctx || !white || !black || !mask) return fail(ENBL_E_ARG, "enbl_nft_mask: null argument");
    if (w < 1 || h < 1 || wpitch < w || bpitch < w || mpitch < w) return fail(ENBL_E_SIZE, "enbl_nft_mask: bad size or pitch");
    if (norm != 1 && norm != 2) return fail(ENBL_E_UNSUPPORTED, "enbl_nft_mask: norm %d (1 Manhattan, 2 Euclidean)", norm);
    const size_t plane = (size_t)w * h;
    void* dbuf;
    int rc = scratch(ctx, 19, 2 * plane * sizeof(float), &dbuf);
    if (rc) return rc;
    NftArgs a;
    memset(&a, 0, sizeof a);
    a.white = white; a.wpitch = wpitch; a.black = black; a.bpitch = bpitch;
    a.w = w; a.h = h; a.norm = norm;
    a.dist = (float*)dbuf; a.dplane = (long)plane; a.dpitch = w;
    a.mask = mask; a.mpitch = mpitch;
    rc = nft_passes(ctx, a, 2);
    if (rc) return rc;
    {
        Timed t(ctx, "nft_mask", (double)plane * 9.0);
        hipLaunchKernelGGL(nft_mask_kernel, dim3((w + 255) / 256, h), dim3(256), 0, ctx->stream, a);
        rc = launch_status("nft_mask");
        if (rc) return rc;
    }
    unsigned tally[2];
    rc = hip_check(hipMemcpyAsync(tally, a.tally, sizeof tally, hipMemcpyDeviceToHost, ctx->stream), "nft tally");
    if (!rc) rc = hip_check(hipStreamSynchronize(ctx->stream), "nft tally");
    if (rc) return rc;
    // too few exclusive pixels in both images: the overlap is (nearly)
    // complete and the seam line is meaningless (src/nearest.h:449-457, :523-538)
    const unsigned long threshold = 2UL * 2UL * ((unsigned long)h + (unsigned long)w);
    if (std::max(tally[0], tally[1]) < threshold)
        return fail(ENBL_E_PRECONDITION, "excessive image overlap detected; too high risk of defective seam line");
    return ENBL_OK;
}

int enbl_nft_mask_periodic(enbl_context* ctx, int w, int h, const uint8_t* white, long wpitch, const uint8_t* black,
                           long bpitch, int norm, int boundary, uint8_t* mask, long mpitch)
{
    if (!ctx || !white || !black || !mask) return fail(ENBL_E_ARG, "enbl_nft_mask_periodic: null argument");
    if (w < 1 || h < 1 || wpitch < w || bpitch < w || mpitch < w)
        return fail(ENBL_E_SIZE, "enbl_nft_mask_periodic: bad size or pitch");
    if (norm != 1 && norm != 2) return fail(ENBL_E_UNSUPPORTED, "enbl_nft_mask_periodic: norm %d", norm);
    if (boundary < 1 || boundary > 3) return fail(ENBL_E_ARG, "enbl_nft_mask_periodic: boundary %d", boundary);
    const int hx = boundary == 1 || boundary == 3, hy = boundary == 2 || boundary == 3;
    const int W = hx ? 2 * w : w, H = hy ? 2 * h : h;
    if ((long)W * H >= (1L << 31)) return fail(ENBL_E_SIZE, "enbl_nft_mask_periodic: image too large");
    const size_t plane = (size_t)W * H;
    void *dbuf, *tiles;
    int rc = scratch(ctx, 19, 2 * plane * sizeof(float), &dbuf);
    if (!rc) rc = scratch(ctx, 31, 2 * plane, &tiles);
    if (rc) return rc;
    uint8_t* w2 = (uint8_t*)tiles;
    uint8_t* b2 = w2 + plane;
    {
        Timed t(ctx, "nft_tile", 2.0 * (double)plane * 2.0);
        const dim3 g((unsigned)((W + 255) / 256), (unsigned)H);
        hipLaunchKernelGGL(nft_tile_kernel, g, dim3(256), 0, ctx->stream, white, wpitch, w, h, w2, W);
        hipLaunchKernelGGL(nft_tile_kernel, g, dim3(256), 0, ctx->stream, black, bpitch, w, h, b2, W);
        rc = launch_status("nft_tile");
        if (rc) return rc;
    }
    NftArgs a;
    memset(&a, 0, sizeof a);
    a.white = w2; a.wpitch = W; a.black = b2; a.bpitch = W;
    a.w = W; a.h = H; a.norm = norm;
    a.dist = (float*)dbuf; a.dplane = (long)plane; a.dpitch = W;
    a.mask = mask; a.mpitch = mpitch;
    rc = nft_passes(ctx, a, 2);
    if (rc) return rc;
    {
        Timed t(ctx, "nft_mask", (double)w * h * 9.0);
        hipLaunchKernelGGL(nft_mask_periodic_kernel, dim3((w + 255) / 256, h), dim3(256), 0, ctx->stream, a, w, h, hx, hy);
        rc = launch_status("nft_mask");
        if (rc) return rc;
    }
    unsigned tally[2];
    rc = hip_check(hipMemcpyAsync(tally, a.tally, sizeof tally, hipMemcpyDeviceToHost, ctx->stream), "nft tally");
    if (!rc) rc = hip_check(hipStreamSynchronize(ctx->stream), "nft tally");
    if (rc) return rc;
    // the tiles hold every exclusive pixel (1 + hx) * (1 + hy) times; the
    // reference tallies the original difference images
    const unsigned copies = (unsigned)((1 + hx) * (1 + hy));
    const unsigned long threshold = 2UL * 2UL * ((unsigned long)h + (unsigned long)w);
    if (std::max(tally[0] / copies, tally[1] / copies) < threshold)
        return fail(ENBL_E_PRECONDITION, "excessive image overlap detected; too high risk of defective seam line");
    return ENBL_OK;
}

}  // extern "C"
