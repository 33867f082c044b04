// enblend's driver around the blend, on the device: import of decoded layers
// with alpha thresholding onto the union canvas (src/assemble.h:231-444), the
// overlap inspection (src/bounds.h:41-80), the seam-mask transition box
// (src/mask.h:520-590), one iteration of enblendMain's blending loop
// (src/enblend.h:127-545) for the fine NFT mask, and the checkpoint
// conversion handed to the exporter (src/assemble.h:130-224).
//
// Every kernel here is a single pass over bytes already in HBM: one thread
// per pixel, consecutive threads on consecutive pixels of a row, so a wave
// reads and writes contiguous runs.  Reductions (bounding boxes, overlap
// flags) fold within the block and leave one set of atomics per block.  The
// host makes the same decisions the reference makes between the passes
// (which layer goes next, which branch of the loop), so those entry points
// read a few bytes back and synchronise the stream.
#include <limits.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "enbl_device.h"
#include "enbl_internal.h"

namespace enbl {
namespace {

constexpr int kDT = 256;        // threads per block, one pixel each
constexpr int kBoxRows = 16;    // rows per block of the reduction kernels
constexpr int kResultInts = 8;  // box (min x, min y, max x1, max y1), two flags, spare

// NumericTraits<T>::fromRealPromote for the unsigned targets: negative and
// NaN to 0, beyond the range to the maximum, else v + 0.5 truncated.
template <class D> __device__ __forceinline__ D from_real_u(double v);
template <> __device__ __forceinline__ uint8_t from_real_u<uint8_t>(double v) {
    return v != v || v < 0.0 ? (uint8_t)0 : v > 255.0 ? (uint8_t)255 : (uint8_t)(int)(v + 0.5);
}
template <> __device__ __forceinline__ uint16_t from_real_u<uint16_t>(double v) {
    return v != v || v < 0.0 ? (uint16_t)0 : v > 65535.0 ? (uint16_t)65535 : (uint16_t)(int)(v + 0.5);
}
template <> __device__ __forceinline__ uint32_t from_real_u<uint32_t>(double v) {
    return v != v || v < 0.0 ? 0u : v > 4294967295.0 ? 4294967295u : (uint32_t)(long long)(v + 0.5);
}
template <> __device__ __forceinline__ float from_real_u<float>(double v) { return (float)v; }

__global__ void init_result_kernel(int* r)
{
    const int t = threadIdx.x;
    if (t < kResultInts) r[t] = t < 2 ? INT_MAX : t < 4 ? INT_MIN : 0;
}

struct Box { int x0, y0, x1, y1; };

__device__ __forceinline__ void box_add(Box& b, int x0, int y0, int x1, int y1)
{
    b.x0 = min(b.x0, x0); b.y0 = min(b.y0, y0);
    b.x1 = max(b.x1, x1); b.y1 = max(b.y1, y1);
}

// Folds the block's boxes and merges the result into r[0..3].
__device__ void block_box(Box b, int* r)
{
    __shared__ int sb[4][kDT / kWave];
    for (int o = kWave / 2; o > 0; o >>= 1) {
        b.x0 = min(b.x0, __shfl_xor(b.x0, o));
        b.y0 = min(b.y0, __shfl_xor(b.y0, o));
        b.x1 = max(b.x1, __shfl_xor(b.x1, o));
        b.y1 = max(b.y1, __shfl_xor(b.y1, o));
    }
    const int wv = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) { sb[0][wv] = b.x0; sb[1][wv] = b.y0; sb[2][wv] = b.x1; sb[3][wv] = b.y1; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kDT / kWave; ++k) box_add(b, sb[0][k], sb[1][k], sb[2][k], sb[3][k]);
        if (b.x0 != INT_MAX) {
            atomicMin(r + 0, b.x0); atomicMin(r + 1, b.y0);
            atomicMax(r + 2, b.x1); atomicMax(r + 3, b.y1);
        }
    }
}

// FindBoundingRectangle over the non-zero pixels of a w x h byte plane.
// With 16-byte aligned rows a thread reads 16 columns per row in one access
// (kBoxRows rows per block); otherwise one column.
template <bool WIDE>
__global__ __launch_bounds__(kDT) void alpha_box_kernel(const uint8_t* __restrict__ a, long ap, int w, int h, int* r)
{
    const int span = WIDE ? 16 : 1;
    const int x = (blockIdx.x * kDT + threadIdx.x) * span;
    const int y0 = blockIdx.y * kBoxRows;
    Box b = {INT_MAX, INT_MAX, INT_MIN, INT_MIN};
    if (x < w) {
        const int ye = min(h, y0 + kBoxRows);
        for (int y = y0; y < ye; ++y) {
            const uint8_t* row = a + (long)y * ap + x;
            if (WIDE && x + 16 <= w) {
                const v4i32 v = *gmem(reinterpret_cast<const v4i32*>(row));
                const uint32_t wd[4] = {(uint32_t)v.x, (uint32_t)v.y, (uint32_t)v.z, (uint32_t)v.w};
                int lo = -1, hi = -1;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    // bytes of word k that are non-zero, as bits 0..3
                    const uint32_t t = wd[k];
                    const uint32_t nz = ((t & 0xffu) != 0) | (((t >> 8) & 0xffu) != 0) << 1 |
                                        (((t >> 16) & 0xffu) != 0) << 2 | ((t >> 24) != 0) << 3;
                    if (nz) {
                        if (lo < 0) lo = 4 * k + __builtin_ctz(nz);
                        hi = 4 * k + 31 - __builtin_clz(nz);
                    }
                }
                if (lo >= 0) box_add(b, x + lo, y, x + hi + 1, y + 1);
            } else {
                const int xe = min(w, x + span);
                for (int xx = x; xx < xe; ++xx)
                    if (*gmem(row + (xx - x))) box_add(b, xx, y, xx + 1, y + 1);
            }
        }
    }
    block_box(b, r);
}

// maskBounds' transition pixels: between (x-1, y) and (x, y) the rectangle
// [x-1, x+1) x [y, y+1); between (x, y-1) and (x, y) the rectangle
// [x, x+1) x [y-1, y+1).
__global__ __launch_bounds__(kDT) void mask_box_kernel(const uint8_t* __restrict__ m, long mp, int w, int h, int* r)
{
    const int x = blockIdx.x * kDT + threadIdx.x;
    const int y0 = blockIdx.y * kBoxRows;
    Box b = {INT_MAX, INT_MAX, INT_MIN, INT_MIN};
    if (x < w) {
        const int ye = min(h, y0 + kBoxRows);
        int up = y0 > 0 ? *gmem(m + (long)(y0 - 1) * mp + x) : 0;
        for (int y = y0; y < ye; ++y) {
            const uint8_t* row = m + (long)y * mp;
            const int v = *gmem(row + x);
            if (x > 0 && *gmem(row + x - 1) != v) box_add(b, x - 1, y, x + 1, y + 1);
            if (y > 0 && up != v) box_add(b, x, y - 1, x + 1, y + 1);
            up = v;
        }
    }
    block_box(b, r);
}

// inspectOverlap's two findings over w x h windows: r[4] = some pixel in
// both, r[5] = some pixel in the second only.
__global__ __launch_bounds__(kDT) void overlap_kernel(const uint8_t* __restrict__ a1, long p1,
                                                      const uint8_t* __restrict__ a2, long p2, int w, int h, int* r)
{
    const int x = blockIdx.x * kDT + threadIdx.x;
    const int y0 = blockIdx.y * kBoxRows;
    int both = 0, only2 = 0;
    if (x < w) {
        const int ye = min(h, y0 + kBoxRows);
        for (int y = y0; y < ye; ++y) {
            const bool s1 = *gmem(a1 + (long)y * p1 + x) != 0, s2 = *gmem(a2 + (long)y * p2 + x) != 0;
            both |= s1 && s2;
            only2 |= s2 && !s1;
        }
    }
    both = __syncthreads_or(both);
    only2 = __syncthreads_or(only2);
    if (threadIdx.x == 0) {
        if (both) *gmem(r + 4) = 1;
        if (only2) *gmem(r + 5) = 1;
    }
}

struct ImportArgs {
    const void* src; long spitch;   // layer samples, elements per row
    int lw, lh, lx, ly, bands;      // layer size, canvas position, bands per pixel
    double lo, hi;                  // alpha threshold bounds, exact in double
    double dscale, doffset;         // linearRangeMapping, double multiplier
    float fscale, foffset;          // linearRangeMapping, float multiplier
    void* img; long ipitch;         // canvas
    uint8_t* alpha; long apitch;
    int uw, uh;
    int* r;                         // overlap flag (overlap pass)
};

// import() of one layer: image bands cast to T, optionally range-mapped
// (MAP 1: double multiplier, MAP 2: float multiplier), alpha band through the
// threshold.  FULL: the grid covers the whole canvas, pixels outside the
// layer become zero (the first layer on a fresh canvas); otherwise the grid
// covers the layer and only pixels with alpha set are written (copyImageIf).
template <class S, class T, int C, int MAP, bool FULL>
__global__ __launch_bounds__(kDT) void import_kernel(ImportArgs a)
{
    const int gx = blockIdx.x * kDT + threadIdx.x, gy = blockIdx.y;
    const int X = FULL ? gx : gx + a.lx, Y = FULL ? gy : gy + a.ly;
    if (FULL ? X >= a.uw : gx >= a.lw) return;
    const int lx = X - a.lx, ly = Y - a.ly;
    const bool in = lx >= 0 && lx < a.lw && ly >= 0 && ly < a.lh;
    T out[C];
    uint8_t al = 0;
#pragma unroll
    for (int c = 0; c < C; ++c) out[c] = (T)0;
    if (in) {
        const S* p = reinterpret_cast<const S*>(a.src) + (long)ly * a.spitch + (long)lx * a.bands;
        al = 255;
        if (a.bands > C) {
            const double d = (double)*gmem(p + C);
            al = (d < a.lo || a.hi < d) ? 0 : 255;
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const T v = (T)*gmem(p + c);
            if constexpr (MAP == 1) out[c] = from_real_u<T>(a.dscale * ((double)v + a.doffset));
            else if constexpr (MAP == 2) out[c] = (T)(a.fscale * ((float)v + a.foffset));
            else out[c] = v;
        }
    }
    if (!FULL && !al) return;
    T* q = reinterpret_cast<T*>(a.img) + (long)Y * a.ipitch + (long)X * C;
#pragma unroll
    for (int c = 0; c < C; ++c) *gmem(q + c) = out[c];
    *gmem(a.alpha + (long)Y * a.apitch + X) = al;
}

// Whether the layer's thresholded alpha meets the canvas alpha anywhere.
template <class S>
__global__ __launch_bounds__(kDT) void layer_overlap_kernel(ImportArgs a, int C)
{
    const int x = blockIdx.x * kDT + threadIdx.x, y = blockIdx.y;
    int hit = 0;
    if (x < a.lw) {
        bool set = true;
        if (a.bands > C) {
            const double d = (double)*gmem(reinterpret_cast<const S*>(a.src) + (long)y * a.spitch + (long)x * a.bands + C);
            set = !(d < a.lo || a.hi < d);
        }
        hit = set && *gmem(a.alpha + (long)(a.ly + y) * a.apitch + (a.lx + x)) != 0;
    }
    hit = __syncthreads_or(hit);
    if (threadIdx.x == 0 && hit) *gmem(a.r + 4) = 1;
}

// copyImageIf of image and alpha under the source alpha (images that do not
// overlap are combined without blending).
template <class T, int C>
__global__ __launch_bounds__(kDT) void copy_if_kernel(const T* __restrict__ s, long sp, const uint8_t* __restrict__ sa,
                                                      long sap, T* d, long dp, uint8_t* da, long dap, int w)
{
    const int x = blockIdx.x * kDT + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const uint8_t av = *gmem(sa + (long)y * sap + x);
    if (!av) return;
#pragma unroll
    for (int c = 0; c < C; ++c) *gmem(d + (long)y * dp + (long)x * C + c) = *gmem(s + (long)y * sp + (long)x * C + c);
    *gmem(da + (long)y * dap + x) = av;
}

// checkpoint(): image samples (mapped when MAP) and the thresholded alpha,
// interleaved as channels + 1 bands of D.
template <class T, class D, int C, bool MAP>
__global__ __launch_bounds__(kDT) void checkpoint_kernel(const T* __restrict__ s, long sp, const uint8_t* __restrict__ sa,
                                                         long sap, D* o, long op, int w, double scale, double offset,
                                                         D amax)
{
    const int x = blockIdx.x * kDT + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const T* p = s + (long)y * sp + (long)x * C;
    D* q = o + (long)y * op + (long)x * (C + 1);
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const T v = *gmem(p + c);
        if constexpr (MAP) *gmem(q + c) = from_real_u<D>(scale * ((double)v + offset));
        else *gmem(q + c) = (D)v;
    }
    *gmem(q + C) = *gmem(sa + (long)y * sap + x) ? amax : (D)0;
}

dim3 grid_rows(int w, int h) { return dim3((unsigned)((w + kDT - 1) / kDT), (unsigned)h); }
dim3 grid_box(int w, int h) { return dim3((unsigned)((w + kDT - 1) / kDT), (unsigned)((h + kBoxRows - 1) / kBoxRows)); }

}  // namespace
}  // namespace enbl

using namespace enbl;
using namespace enbl::host;

namespace {

// The small result block (box and flags) in context scratch slot 24.
int result_buffer(enbl_context* ctx, int** r)
{
    void* p = nullptr;
    const int rc = scratch(ctx, 24, kResultInts * sizeof(int), &p);
    *r = (int*)p;
    if (rc) return rc;
    hipLaunchKernelGGL(init_result_kernel, dim3(1), dim3(kWave), 0, ctx->stream, *r);
    return launch_status("init_result");
}

int read_result(enbl_context* ctx, const int* r, int* host)
{
    int rc = hip_check(hipMemcpyAsync(host, r, kResultInts * sizeof(int), hipMemcpyDeviceToHost, ctx->stream),
                       "result readback");
    if (!rc) rc = hip_check(hipStreamSynchronize(ctx->stream), "result readback");
    return rc;
}

enbl_rect box_of(const int* h)
{
    // FindBoundingRectangle of no pixel: Rect2D() = (0, 0, 0, 0)
    if (h[0] == INT_MAX) return enbl_rect{0, 0, 0, 0};
    return enbl_rect{h[0], h[1], h[2], h[3]};
}

bool empty(const enbl_rect& r) { return r.x1 <= r.x0 || r.y1 <= r.y0; }

enbl_rect rect_union(enbl_rect a, const enbl_rect& b)
{
    if (empty(b)) return a;
    if (empty(a)) return b;
    return enbl_rect{std::min(a.x0, b.x0), std::min(a.y0, b.y0), std::max(a.x1, b.x1), std::max(a.y1, b.y1)};
}

enbl_rect rect_intersect(enbl_rect a, const enbl_rect& b)
{
    if (empty(a)) return a;
    if (empty(b)) return b;
    return enbl_rect{std::max(a.x0, b.x0), std::max(a.y0, b.y0), std::min(a.x1, b.x1), std::min(a.y1, b.y1)};
}

double range_max(int t) { return t == ENBL_U8 ? 255.0 : t == ENBL_U16 ? 65535.0 : 1.0; }

bool layer_types_ok(int itype, int ftype)
{
    if (ftype == itype) return true;
    return (ftype == ENBL_U8 && (itype == ENBL_U16 || itype == ENBL_F32)) || (ftype == ENBL_U16 && itype == ENBL_F32);
}

// Import parameters of one layer: thresholds T(max/2), T(max) as doubles and
// linearRangeMapping(T(0), T(fmax), T(0), T(tmax)) when the file's range
// differs from the image's.
template <class T>
void import_params(const enbl_layer& L, ImportArgs& a, int* map)
{
    const double fmax = range_max(L.ftype);
    a.lo = (double)(T)(fmax / 2.0);
    a.hi = (double)(T)fmax;
    const double tmax = range_max(sizeof(T) == 1 ? ENBL_U8 : sizeof(T) == 2 ? ENBL_U16 : ENBL_F32);
    *map = fmax != tmax ? (sizeof(T) == 4 ? 2 : 1) : 0;
    if (*map == 1) {
        const double diff = (double)((int)(T)fmax - (int)(T)0.0);
        a.dscale = diff == 0.0 ? 1.0 : (double)((int)(T)tmax - (int)(T)0.0) / diff;
        a.doffset = (double)(T)0.0 / a.dscale - (double)(T)0.0;
    } else if (*map == 2) {
        const float diff = (float)fmax - 0.0f;
        a.fscale = diff == 0.0f ? 1.0f : (1.0f - 0.0f) / diff;
        a.foffset = 0.0f / a.fscale - 0.0f;
    }
}

template <class S, class T, int MAP>
void launch_import_t(const ImportArgs& a, int C, bool full, hipStream_t s)
{
    const dim3 g = full ? grid_rows(a.uw, a.uh) : grid_rows(a.lw, a.lh);
    if (C == 3) {
        if (full) hipLaunchKernelGGL((import_kernel<S, T, 3, MAP, true>), g, dim3(kDT), 0, s, a);
        else hipLaunchKernelGGL((import_kernel<S, T, 3, MAP, false>), g, dim3(kDT), 0, s, a);
    } else {
        if (full) hipLaunchKernelGGL((import_kernel<S, T, 1, MAP, true>), g, dim3(kDT), 0, s, a);
        else hipLaunchKernelGGL((import_kernel<S, T, 1, MAP, false>), g, dim3(kDT), 0, s, a);
    }
}

// The file/image type pairs import() handles: equal types unchanged, 8-bit
// into 16-bit with a double multiplier, 8- or 16-bit into float with a float
// multiplier.
template <class T>
int import_layer_t(enbl_context* ctx, int C, const enbl_layer& L, ImportArgs a, bool full)
{
    int map = 0;
    import_params<T>(L, a, &map);
    hipStream_t s = ctx->stream;
    const double px = full ? (double)a.uw * a.uh : (double)L.w * L.h;
    Timed t(ctx, full ? "import_layer_full" : "import_layer_if",
            px * (L.bands * ptype_size(L.ftype) + C * sizeof(T) + 1.0));
    if (sizeof(T) == 1) launch_import_t<uint8_t, uint8_t, 0>(a, C, full, s);
    else if (sizeof(T) == 2) {
        if (L.ftype == ENBL_U8) launch_import_t<uint8_t, uint16_t, 1>(a, C, full, s);
        else launch_import_t<uint16_t, uint16_t, 0>(a, C, full, s);
    } else {
        if (L.ftype == ENBL_U8) launch_import_t<uint8_t, float, 2>(a, C, full, s);
        else if (L.ftype == ENBL_U16) launch_import_t<uint16_t, float, 2>(a, C, full, s);
        else launch_import_t<float, float, 0>(a, C, full, s);
    }
    return launch_status("import_layer");
}

int import_layer(enbl_context* ctx, int itype, int C, const enbl_layer& L, const ImportArgs& a, bool full)
{
    if (itype == ENBL_U8) return import_layer_t<uint8_t>(ctx, C, L, a, full);
    if (itype == ENBL_U16) return import_layer_t<uint16_t>(ctx, C, L, a, full);
    return import_layer_t<float>(ctx, C, L, a, full);
}

int layer_overlaps(enbl_context* ctx, int itype, int C, const enbl_layer& L, ImportArgs a, bool* hit)
{
    int* r = nullptr;
    int rc = result_buffer(ctx, &r);
    if (rc) return rc;
    a.r = r;
    int map = 0;
    if (itype == ENBL_U8) import_params<uint8_t>(L, a, &map);
    else if (itype == ENBL_U16) import_params<uint16_t>(L, a, &map);
    else import_params<float>(L, a, &map);
    {
        Timed t(ctx, "layer_overlap", (double)L.w * L.h * (ptype_size(L.ftype) + 1.0));
        const dim3 g = grid_rows(L.w, L.h);
        if (L.ftype == ENBL_U8) hipLaunchKernelGGL(layer_overlap_kernel<uint8_t>, g, dim3(kDT), 0, ctx->stream, a, C);
        else if (L.ftype == ENBL_U16) hipLaunchKernelGGL(layer_overlap_kernel<uint16_t>, g, dim3(kDT), 0, ctx->stream, a, C);
        else hipLaunchKernelGGL(layer_overlap_kernel<float>, g, dim3(kDT), 0, ctx->stream, a, C);
        rc = launch_status("layer_overlap");
        if (rc) return rc;
    }
    int h[kResultInts];
    rc = read_result(ctx, r, h);
    *hit = h[4] != 0;
    return rc;
}

int plane_box(enbl_context* ctx, bool transitions, const uint8_t* p, long pitch, int w, int h, int* host)
{
    int* r = nullptr;
    int rc = result_buffer(ctx, &r);
    if (rc) return rc;
    if (w > 0 && h > 0) {
        Timed t(ctx, transitions ? "mask_bounds" : "alpha_box", (double)w * h);
        if (transitions) hipLaunchKernelGGL(mask_box_kernel, grid_box(w, h), dim3(kDT), 0, ctx->stream, p, pitch, w, h, r);
        else if (((uintptr_t)p & 15) == 0 && (pitch & 15) == 0)
            hipLaunchKernelGGL(alpha_box_kernel<true>, grid_box((w + 15) / 16, h), dim3(kDT), 0, ctx->stream, p, pitch, w,
                               h, r);
        else hipLaunchKernelGGL(alpha_box_kernel<false>, grid_box(w, h), dim3(kDT), 0, ctx->stream, p, pitch, w, h, r);
        rc = launch_status(transitions ? "mask_bounds" : "alpha_box");
        if (rc) return rc;
    }
    return read_result(ctx, r, host);
}

int overlap_of(enbl_context* ctx, int w, int h, const uint8_t* a1, long p1, const uint8_t* a2, long p2, int* ov)
{
    int* r = nullptr;
    int rc = result_buffer(ctx, &r);
    if (rc) return rc;
    if (w > 0 && h > 0) {
        Timed t(ctx, "inspect_overlap", 2.0 * w * h);
        hipLaunchKernelGGL(overlap_kernel, grid_box(w, h), dim3(kDT), 0, ctx->stream, a1, p1, a2, p2, w, h, r);
        rc = launch_status("inspect_overlap");
        if (rc) return rc;
    }
    int hv[kResultInts];
    rc = read_result(ctx, r, hv);
    if (rc) return rc;
    // PartialOverlap once both findings exist, CompleteOverlap with overlap
    // alone, NoOverlap otherwise
    *ov = hv[4] ? (hv[5] ? 1 : 2) : 0;
    return 0;
}

template <class T>
int copy_if_t(enbl_context* ctx, int C, int w, int h, const void* s, long sp, const uint8_t* sa, long sap, void* d,
              long dp, uint8_t* da, long dap)
{
    Timed t(ctx, "copy_if", (double)w * h * (2.0 * C * sizeof(T) + 2.0));
    if (C == 3)
        hipLaunchKernelGGL((copy_if_kernel<T, 3>), grid_rows(w, h), dim3(kDT), 0, ctx->stream, (const T*)s, sp, sa, sap,
                           (T*)d, dp, da, dap, w);
    else
        hipLaunchKernelGGL((copy_if_kernel<T, 1>), grid_rows(w, h), dim3(kDT), 0, ctx->stream, (const T*)s, sp, sa, sap,
                           (T*)d, dp, da, dap, w);
    return launch_status("copy_if");
}

template <class T, class D>
int checkpoint_t(enbl_context* ctx, int C, int w, int h, const void* img, long ip, const uint8_t* a, long ap, void* out,
                 long op, bool map, double scale, double offset, D amax)
{
    Timed t(ctx, "checkpoint", (double)w * h * (C * sizeof(T) + 1.0 + (C + 1) * sizeof(D)));
    const dim3 g = grid_rows(w, h);
#define CK(CC, M)                                                                                                   \
    hipLaunchKernelGGL((checkpoint_kernel<T, D, CC, M>), g, dim3(kDT), 0, ctx->stream, (const T*)img, ip, a, ap, \
                       (D*)out, op, w, scale, offset, amax)
    if (C == 3) { if (map) CK(3, true); else CK(3, false); }
    else { if (map) CK(1, true); else CK(1, false); }
#undef CK
    return launch_status("checkpoint");
}

}  // namespace

extern "C" int enbl_assemble(enbl_context* ctx, int itype, int C, int uw, int uh, int n, const enbl_layer* layers,
                             int one_at_a_time, int* taken, void* image, long ipitch, uint8_t* alpha, long apitch,
                             enbl_rect* bb)
{
    if (!ctx || (n > 0 && (!layers || !taken)) || !image || !alpha || !bb)
        return fail(ENBL_E_ARG, "enbl_assemble: null argument");
    if (itype != ENBL_U8 && itype != ENBL_U16 && itype != ENBL_F32) return fail(ENBL_E_UNSUPPORTED, "image type %d", itype);
    if (C != 1 && C != 3) return fail(ENBL_E_UNSUPPORTED, "enbl_assemble: %d channels", C);
    if (uw <= 0 || uh <= 0 || ipitch < (long)uw * C || apitch < uw) return fail(ENBL_E_ARG, "enbl_assemble: canvas size");
    for (int i = 0; i < n; ++i) {
        const enbl_layer& L = layers[i];
        if (!L.data && L.w > 0 && L.h > 0) return fail(ENBL_E_ARG, "enbl_assemble: layer %d has no data", i);
        if (!layer_types_ok(itype, L.ftype))
            return fail(ENBL_E_UNSUPPORTED, "enbl_assemble: layer %d of type %d into image type %d", i, L.ftype, itype);
        if (L.bands != C && L.bands != C + 1) return fail(ENBL_E_ARG, "enbl_assemble: layer %d has %d bands", i, L.bands);
        if (L.w < 0 || L.h < 0 || L.x < 0 || L.y < 0 || L.x > uw - L.w || L.y > uh - L.h || L.pitch < (long)L.w * L.bands)
            return fail(ENBL_E_ARG, "enbl_assemble: layer %d does not fit the canvas", i);
    }
    int first = -1;
    for (int i = 0; i < n; ++i)
        if (!taken[i]) { first = i; break; }
    if (first < 0) return 0;

    ImportArgs a;
    memset(&a, 0, sizeof a);
    a.img = image; a.ipitch = ipitch; a.alpha = alpha; a.apitch = apitch; a.uw = uw; a.uh = uh;
    auto set_layer = [&](const enbl_layer& L) {
        a.src = L.data; a.spitch = L.pitch; a.lw = L.w; a.lh = L.h; a.lx = L.x; a.ly = L.y; a.bands = L.bands;
    };
    set_layer(layers[first]);
    int rc = import_layer(ctx, itype, C, layers[first], a, true);
    if (rc) return rc;
    int placed = 1;
    std::vector<int> accepted;
    if (!one_at_a_time) {
        for (int i = first + 1; i < n; ++i) {
            if (taken[i] || layers[i].w == 0 || layers[i].h == 0) {
                if (!taken[i]) { accepted.push_back(i); ++placed; }   // nothing to import, nothing to overlap
                continue;
            }
            set_layer(layers[i]);
            bool hit = false;
            rc = layer_overlaps(ctx, itype, C, layers[i], a, &hit);
            if (rc) return rc;
            if (hit) continue;
            rc = import_layer(ctx, itype, C, layers[i], a, false);
            if (rc) return rc;
            accepted.push_back(i);
            ++placed;
        }
    }
    taken[first] = 1;
    for (int i : accepted) taken[i] = 1;
    int h[kResultInts];
    rc = plane_box(ctx, false, alpha, apitch, uw, uh, h);
    if (rc) return rc;
    *bb = box_of(h);
    return placed;
}

extern "C" int enbl_inspect_overlap(enbl_context* ctx, int w, int h, const uint8_t* a1, long p1, const uint8_t* a2,
                                    long p2, int* overlap)
{
    if (!ctx || !overlap || ((w > 0 && h > 0) && (!a1 || !a2))) return fail(ENBL_E_ARG, "enbl_inspect_overlap: null argument");
    if (w < 0 || h < 0 || p1 < w || p2 < w) return fail(ENBL_E_ARG, "enbl_inspect_overlap: size");
    return overlap_of(ctx, w, h, a1, p1, a2, p2, overlap);
}

extern "C" int enbl_mask_bounds(enbl_context* ctx, int w, int h, const uint8_t* mask, long mpitch, enbl_rect* bb,
                                int* found)
{
    if (!ctx || !bb || !found || ((w > 0 && h > 0) && !mask)) return fail(ENBL_E_ARG, "enbl_mask_bounds: null argument");
    if (w < 0 || h < 0 || mpitch < w) return fail(ENBL_E_ARG, "enbl_mask_bounds: size");
    int hv[kResultInts];
    const int rc = plane_box(ctx, true, mask, mpitch, w, h, hv);
    if (rc) return rc;
    *found = hv[0] != INT_MAX;
    *bb = *found ? enbl_rect{hv[0], hv[1], hv[2], hv[3]} : enbl_rect{w, h, 0, 0};
    return 0;
}

extern "C" int enbl_enblend_step(enbl_context* ctx, int itype, int C, int uw, int uh, int wrap, int exact_levels,
                                 int minimum_levels, int norm, void* black, long bpitch, uint8_t* balpha, long bapitch,
                                 enbl_rect* black_bb, const void* white, long wpitch, const uint8_t* walpha,
                                 long wapitch, const enbl_rect* white_bb, int dither, uint64_t seed, int* outcome)
{
    if (!ctx || !black || !balpha || !black_bb || !white || !walpha || !white_bb || !outcome)
        return fail(ENBL_E_ARG, "enbl_enblend_step: null argument");
    if (itype != ENBL_U8 && itype != ENBL_U16 && itype != ENBL_F32) return fail(ENBL_E_UNSUPPORTED, "image type %d", itype);
    if (C != 1 && C != 3) return fail(ENBL_E_UNSUPPORTED, "enbl_enblend_step: %d channels", C);
    if (norm != 1 && norm != 2) return fail(ENBL_E_ARG, "enbl_enblend_step: norm %d", norm);
    // the ROI border filterHalfWidth(29) = 2^30 - 4 must not overflow
    if (uw <= 0 || uh <= 0 || uw >= (1 << 29) || uh >= (1 << 29)) return fail(ENBL_E_ARG, "enbl_enblend_step: canvas size");
    if (bpitch < (long)uw * C || wpitch < (long)uw * C || bapitch < uw || wapitch < uw)
        return fail(ENBL_E_ARG, "enbl_enblend_step: pitch");
    const auto inside = [&](const enbl_rect& r) {
        return empty(r) || (r.x0 >= 0 && r.y0 >= 0 && r.x1 <= uw && r.y1 <= uh);
    };
    if (!inside(*black_bb) || !inside(*white_bb)) return fail(ENBL_E_ARG, "enbl_enblend_step: box outside the canvas");
    const size_t esz = (size_t)ptype_size(itype);
    const enbl_rect uBB = rect_union(*black_bb, *white_bb);
    const long ua = (long)uBB.y0 * bapitch + uBB.x0, uwa = (long)uBB.y0 * wapitch + uBB.x0;

    // Overlap of black (first) and white (second) over uBB.
    int ov = 0;
    int rc = 0;
    if (!empty(uBB)) {
        rc = overlap_of(ctx, uBB.x1 - uBB.x0, uBB.y1 - uBB.y0, balpha + ua, bapitch, walpha + uwa, wapitch, &ov);
        if (rc) return rc;
    }
    if (ov == 2) {   // CompleteOverlap: the white image is redundant
        *outcome = 1;
        return 0;
    }
    if (ov == 0 && exact_levels == 0) {   // NoOverlap: white copied into black verbatim
        if (itype == ENBL_U8) rc = copy_if_t<uint8_t>(ctx, C, uw, uh, white, wpitch, walpha, wapitch, black, bpitch, balpha, bapitch);
        else if (itype == ENBL_U16) rc = copy_if_t<uint16_t>(ctx, C, uw, uh, white, wpitch, walpha, wapitch, black, bpitch, balpha, bapitch);
        else rc = copy_if_t<float>(ctx, C, uw, uh, white, wpitch, walpha, wapitch, black, bpitch, balpha, bapitch);
        if (rc) return rc;
        *black_bb = uBB;
        *outcome = 2;
        return 0;
    }
    const int w = uBB.x1 - uBB.x0, h = uBB.y1 - uBB.y0;
    const bool wrap_mask = wrap && w == uw;   // wraparoundForMask
    if (empty(uBB)) return fail(ENBL_E_PRECONDITION, "enbl_enblend_step: neither image has an alpha pixel");

    // Seam mask over uBB: nearest-feature transform of the two alphas.
    void* mp = nullptr;
    rc = scratch(ctx, 25, (size_t)w * h, &mp);
    if (rc) return rc;
    uint8_t* mask = (uint8_t*)mp;
    rc = wrap_mask ? enbl_nft_mask_periodic(ctx, w, h, walpha + uwa, wapitch, balpha + ua, bapitch, norm, 1, mask, w)
                   : enbl_nft_mask(ctx, w, h, walpha + uwa, wapitch, balpha + ua, bapitch, norm, mask, w);
    if (rc) return rc;

    // maskBounds: mBB relative to the canvas.
    enbl_rect mBB;
    int found = 0;
    rc = enbl_mask_bounds(ctx, w, h, mask, w, &mBB, &found);
    if (rc) return rc;
    if (!found) {
        uint8_t m0 = 0;
        rc = hip_check(hipMemcpyAsync(&m0, mask, 1, hipMemcpyDeviceToHost, ctx->stream), "mask readback");
        if (!rc) rc = hip_check(hipStreamSynchronize(ctx->stream), "mask readback");
        if (rc) return rc;
        if (m0 == 0)
            return fail(ENBL_E_PRECONDITION, "mask is entirely black, but white image was not identified as redundant");
        mBB = uBB;
    } else {
        mBB = enbl_rect{mBB.x0 + uBB.x0, mBB.y0 + uBB.y0, mBB.x1 + uBB.x0, mBB.y1 + uBB.y0};
    }

    // roiBounds: mBB grown by filterHalfWidth(29) = 2^30 - 4 (made [0, uBB.x1)
    // wide when the mask wraps and the box leaves that) and clipped to uBB,
    // which is all of uBB for any canvas below 2^29 pixels a side; the
    // reference's copy of white pixels outside the ROI and its alpha union
    // outside the ROI are then empty.
    const int border = (1 << 30) - 4;
    enbl_rect grown{mBB.x0 - border, mBB.y0 - border, mBB.x1 + border, mBB.y1 + border};
    if (wrap_mask && (grown.x0 < 0 || grown.x1 > uBB.x1)) { grown.x0 = 0; grown.x1 = uBB.x1; }
    const enbl_rect roi = rect_intersect(grown, uBB);
    if (roi.x0 != uBB.x0 || roi.y0 != uBB.y0 || roi.x1 != uBB.x1 || roi.y1 != uBB.y1)
        return fail(ENBL_E_UNSUPPORTED, "enbl_enblend_step: ROI smaller than the union box");
    const int levels = enbl_auto_levels(w, h, minimum_levels, exact_levels);
    const int wrap_blend = wrap && w == uw;
    const char* wb = (const char*)white + ((size_t)uBB.y0 * wpitch + (size_t)uBB.x0 * C) * esz;
    char* bk = (char*)black + ((size_t)uBB.y0 * bpitch + (size_t)uBB.x0 * C) * esz;
    rc = enbl_enblend_pair(ctx, itype, C, w, h, levels, wrap_blend, wb, wpitch, walpha + uwa, wapitch, bk, bpitch,
                           balpha + ua, bapitch, mask, w, dither, seed);
    if (rc) return rc;
    *black_bb = uBB;
    *outcome = 0;
    return 0;
}

extern "C" int enbl_checkpoint(enbl_context* ctx, int itype, int C, int w, int h, const void* img, long ipitch,
                               const uint8_t* alpha, long apitch, int otype, void* out, long opitch, int* out_ctype)
{
    if (!ctx || !out_ctype || ((w > 0 && h > 0) && (!img || !alpha || !out)))
        return fail(ENBL_E_ARG, "enbl_checkpoint: null argument");
    if (itype != ENBL_U8 && itype != ENBL_U16 && itype != ENBL_F32) return fail(ENBL_E_UNSUPPORTED, "image type %d", itype);
    if (otype != ENBL_U8 && otype != ENBL_U16 && otype != ENBL_F32) return fail(ENBL_E_UNSUPPORTED, "output type %d", otype);
    if (C != 1 && C != 3) return fail(ENBL_E_UNSUPPORTED, "enbl_checkpoint: %d channels", C);
    if (w < 0 || h < 0 || ipitch < (long)w * C || apitch < w || opitch < (long)w * (C + 1))
        return fail(ENBL_E_ARG, "enbl_checkpoint: size");
    const double imin = 0.0, imax = range_max(itype), omin = 0.0, omax = range_max(otype);
    if (imin <= omin && imax >= omax) {
        // the image's range covers the output's: unchanged when equal, else
        // narrowed by linearRangeMapping within the image's own type
        const bool map = !(imin == omin && imax == omax);
        *out_ctype = itype;
        if (w == 0 || h == 0) return 0;
        if (itype == ENBL_U8) {
            const double diff = (double)((int)(uint8_t)imax - (int)(uint8_t)imin);
            const double scale = diff == 0.0 ? 1.0 : (double)((int)(uint8_t)omax - (int)(uint8_t)omin) / diff;
            const double offset = (double)(uint8_t)omin / scale - (double)(uint8_t)imin;
            return checkpoint_t<uint8_t, uint8_t>(ctx, C, w, h, img, ipitch, alpha, apitch, out, opitch, map, scale,
                                                  offset, (uint8_t)255);
        }
        if (itype == ENBL_U16) {
            const double diff = (double)((int)(uint16_t)imax - (int)(uint16_t)imin);
            const double scale = diff == 0.0 ? 1.0 : (double)((int)(uint16_t)omax - (int)(uint16_t)omin) / diff;
            const double offset = (double)(uint16_t)omin / scale - (double)(uint16_t)imin;
            return checkpoint_t<uint16_t, uint16_t>(ctx, C, w, h, img, ipitch, alpha, apitch, out, opitch, map, scale,
                                                    offset, (uint16_t)65535);
        }
        return checkpoint_t<float, float>(ctx, C, w, h, img, ipitch, alpha, apitch, out, opitch, false, 1.0, 0.0, 1.0f);
    }
    if (itype != ENBL_F32)
        return fail(ENBL_E_UNSUPPORTED, "enbl_checkpoint: image type %d written as %d overflows the conversion", itype, otype);
    // a float image rescaled into IntegralSelect<RGBValue<float>> = UInt32
    *out_ctype = ENBL_U32;
    if (w == 0 || h == 0) return 0;
    const double diff = (double)((float)imax - (float)imin);
    const double scale = diff == 0.0 ? 1.0 : (double)((uint32_t)omax - (uint32_t)omin) / diff;
    const double offset = (double)(uint32_t)omin / scale - (double)(float)imin;
    return checkpoint_t<float, uint32_t>(ctx, C, w, h, img, ipitch, alpha, apitch, out, opitch, true, scale, offset,
                                         4294967295u);
}

extern "C" int enbl_enblend_layers(enbl_context* ctx, int itype, int C, int uw, int uh, int n,
                                   const enbl_layer* layers, int pre_assemble, int wrap, int exact_levels,
                                   int minimum_levels, int norm, int dither, uint64_t seed, void* black, long bpitch,
                                   uint8_t* balpha, long bapitch, enbl_rect* black_bb, int* counts)
{
    if (!ctx || (n > 0 && !layers) || !black || !balpha || !black_bb)
        return fail(ENBL_E_ARG, "enbl_enblend_layers: null argument");
    if (itype != ENBL_U8 && itype != ENBL_U16 && itype != ENBL_F32) return fail(ENBL_E_UNSUPPORTED, "image type %d", itype);
    if (C != 1 && C != 3) return fail(ENBL_E_UNSUPPORTED, "enbl_enblend_layers: %d channels", C);
    if (n <= 0) return fail(ENBL_E_ARG, "enbl_enblend_layers: no input image");
    if (uw <= 0 || uh <= 0) return fail(ENBL_E_ARG, "enbl_enblend_layers: canvas size");
    std::vector<int> taken((size_t)n, 0);
    int tally[3] = {0, 0, 0};   // blended, redundant, copied without blending
    // the initial black image
    int rc = enbl_assemble(ctx, itype, C, uw, uh, n, layers, !pre_assemble, taken.data(), black, bpitch, balpha,
                           bapitch, black_bb);
    if (rc < 0) return rc;
    // the white canvas: context scratch, packed rows
    const size_t esz = (size_t)ptype_size(itype);
    void *wimg = nullptr, *wal = nullptr;
    rc = scratch(ctx, 22, (size_t)uw * uh * C * esz, &wimg);
    if (!rc) rc = scratch(ctx, 26, (size_t)uw * uh, &wal);
    if (rc) return rc;
    for (int m = 0;; ++m) {
        bool left = false;
        for (int i = 0; i < n; ++i) left |= !taken[i];
        if (!left) break;
        enbl_rect wbb;
        rc = enbl_assemble(ctx, itype, C, uw, uh, n, layers, !pre_assemble, taken.data(), wimg, (long)uw * C,
                           (uint8_t*)wal, uw, &wbb);
        if (rc < 0) return rc;
        int outcome = 0;
        rc = enbl_enblend_step(ctx, itype, C, uw, uh, wrap, exact_levels, minimum_levels, norm, black, bpitch, balpha,
                               bapitch, black_bb, wimg, (long)uw * C, (const uint8_t*)wal, uw, &wbb, dither,
                               seed + (uint64_t)m, &outcome);
        if (rc) return rc;
        ++tally[outcome];
    }
    if (counts) { counts[0] = tally[0]; counts[1] = tally[1]; counts[2] = tally[2]; }
    return hip_check(hipStreamSynchronize(ctx->stream), "enbl_enblend_layers");
}
