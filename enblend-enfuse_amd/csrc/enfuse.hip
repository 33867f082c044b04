// enfuse weighting, normalisation and level merge for gfx950.
//
// Reference: src/enfuse.h:64-1212 (enfuseMask and its functors) and
// :1497-1601 (normalisation, mask pyramid, multiply, accumulate).
//
// * Integer images never evaluate the exposure weight on the device: the host
//   tabulates (float)(W * weight(code / max)) for every gray code, so the
//   user ExposureWeight interface and the built-ins share one exact path.
// * The local standard deviation of integer images uses exact integer window
//   sums, so any summation order reproduces the reference.  Float images keep
//   the reference's float running sums: one thread walks one row in the
//   reference's order.
// * The local entropy sorts each channel's window samples and sums the
//   host-tabulated p*log(p) terms in ascending key order, which is the order
//   of the reference's std::map histograms.
#include <float.h>
#include <string.h>

#include "enbl_device.h"
#include "enbl_enfuse.h"
#include "enbl_internal.h"

namespace enbl {
namespace {

constexpr int kTile = 256;

inline dim3 grid_for(int w, int h) { return dim3((w + kTile - 1) / kTile, h); }

template <class T> struct Img;
template <> struct Img<uint8_t> { static constexpr bool integral = true; __device__ static double maxv() { return 255.0; } };
template <> struct Img<uint16_t> { static constexpr bool integral = true; __device__ static double maxv() { return 65535.0; } };
template <> struct Img<float> { static constexpr bool integral = false; __device__ static double maxv() { return (double)FLT_MAX; } };

template <class T> __device__ T from_real_img(double v);
template <> __device__ __forceinline__ uint8_t from_real_img<uint8_t>(double v) {
    if (v != v || v < 0.0) return 0;
    return v > 255.0 ? (uint8_t)255 : (uint8_t)(int32_t)(v + 0.5);
}
template <> __device__ __forceinline__ uint16_t from_real_img<uint16_t>(double v) {
    if (v != v || v < 0.0) return 0;
    return v > 65535.0 ? (uint16_t)65535 : (uint16_t)(int32_t)(v + 0.5);
}
template <> __device__ __forceinline__ float from_real_img<float>(double v) { return (float)v; }

// "average" grayscale projector (src/mga.h:188-193).
// Grayscale projectors (MultiGrayscaleAccessor, src/mga.h:60-250).  Integer
// channels promote to double, float channels to float; every result lies in
// the channel range, so rounding it to the channel type and widening it to R
// equals rounding it to R.  l-star / pl-star: max * L* / 100 of VIGRA's
// RGB2LabFunctor / RGBPrime2LabFunctor (Y of the sRGB matrix; gamma 1/0.45
// for the primed one; L* = Y < 216/24389 ? 24389/27 Y : 116 Y^(1/3) - 16).
template <class T> struct RealOf { typedef double R; };
template <> struct RealOf<float> { typedef float R; };
template <class T> __device__ __forceinline__ double chan_max();
template <> __device__ __forceinline__ double chan_max<uint8_t>() { return 255.0; }
template <> __device__ __forceinline__ double chan_max<uint16_t>() { return 65535.0; }
template <> __device__ __forceinline__ double chan_max<float>() { return (double)FLT_MAX; }

__device__ __forceinline__ double lstar(double r, double g, double b)
{
    const double Y = 0.212671 * r + 0.715160 * g + 0.072169 * b;
    return Y < 216.0 / 24389.0 ? 24389.0 / 27.0 * Y : 116.0 * pow(Y, 1.0 / 3.0) - 16.0;
}
__device__ __forceinline__ double gamma_correct(double v, double gm) { return v < 0.0 ? -pow(-v, gm) : pow(v, gm); }

template <class T>
__device__ __forceinline__ T project(const T* p, const GrayProj& pj)
{
    typedef typename RealOf<T>::R RP;
    const T r = p[0], g = p[1], b = p[2];
    switch (pj.kind) {
    case ENBL_GRAY_LSTAR:
    case ENBL_GRAY_PLSTAR: {
        const double mx = chan_max<T>();
        double rr = (double)r / mx, gg = (double)g / mx, bb = (double)b / mx;
        if (pj.kind == ENBL_GRAY_PLSTAR) {
            rr = gamma_correct(rr, 1.0 / 0.45);
            gg = gamma_correct(gg, 1.0 / 0.45);
            bb = gamma_correct(bb, 1.0 / 0.45);
        }
        return from_real_img<T>(mx * (lstar(rr, gg, bb) / 100.0));
    }
    case ENBL_GRAY_LIGHTNESS: {
        const T mn = min(r, min(g, b)), mx = max(r, max(g, b));
        return from_real_img<T>((double)(mn + mx) / 2.0);
    }
    case ENBL_GRAY_VALUE: return max(r, max(g, b));
    case ENBL_GRAY_ANTI_VALUE: return min(r, min(g, b));
    case ENBL_GRAY_LUMINANCE: return from_real_img<T>(0.3 * (double)r + 0.59 * (double)g + 0.11 * (double)b);
    case ENBL_GRAY_MIXER:
        return from_real_img<T>(pj.wr * (double)(RP)r + pj.wg * (double)(RP)g + pj.wb * (double)(RP)b);
    default:
        return from_real_img<T>((double)((RP)r + (RP)g + (RP)b) / 3.0);
    }
}

template <class T, class R> struct Gray {
    __device__ static R get(const T* p, const GrayProj& pj) { return (R)project<T>(p, pj); }
};

__device__ double exposure_builtin(int fn, double opt, double width, double y)
{
    const double z = (y - opt) / width;
    switch (fn) {
    case ENBL_EXPOSURE_GAUSSIAN: return exp(-0.5 * z * z);
    case ENBL_EXPOSURE_LORENTZIAN: return 1.0 / (1.0 + 0.5 * z * z);
    case ENBL_EXPOSURE_HALFSINE: return fabs(z) <= M_PI_2 ? cos(z) : 0.0;
    case ENBL_EXPOSURE_FULLSINE: return fabs(z) <= M_PI ? 0.5 + cos(z) / 2.0 : 0.0;
    default:
        if (fabs(z) <= 1.0) { const double z2 = z * z; return (1.0 - z2) * (1.0 + z2); }
        return 0.0;
    }
}

// Gray image for the contrast weight: int for integer images, float for
// float images (LongScalarType = Promote<ScalarType>).
template <class T, class G>
__global__ __launch_bounds__(kTile) void gray_kernel(const T* __restrict__ img, long ip, int w, G* __restrict__ gray, long gp,
                                                     GrayProj pj)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    gray[(long)y * gp + x] = Gray<T, G>::get(img + (long)y * ip + 3L * x, pj);
}

// localStdDevIf for integer gray images (src/enfuse.h:75-219): exact window
// sums, then the reference's double formula; interior pixels under alpha.
__global__ __launch_bounds__(kTile) void stddev_int_kernel(const int32_t* __restrict__ gray, long gp,
                                                           const uint8_t* __restrict__ alpha, long ap,
                                                           int w, int h, int n, int32_t* __restrict__ out, long op)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const int b = n / 2;
    int32_t r = 0;
    if (x >= b && x < w - b && y >= b && y < h - b && alpha[(long)y * ap + x]) {
        long long s = 0, q = 0;
        int cnt = 0;
        for (int yy = y - b; yy <= y + b; ++yy)
            for (int xx = x - b; xx <= x + b; ++xx)
                if (alpha[(long)yy * ap + xx]) {
                    const long long v = gray[(long)yy * gp + xx];
                    s += v;
                    q += v * v;
                    ++cnt;
                }
        if (cnt > 1) {
            const double ds = (double)s, dq = (double)q;
            r = from_real<int32_t>(sqrt((dq - ds * ds / (double)cnt) / (double)(cnt - 1)));
        }
    }
    out[(long)y * op + x] = r;
}

// localStdDevIf for float gray images: one thread per row, float running
// sums in the reference's order (column sums top to bottom, window total
// updated by (new - old) per step).
template <int NMAX>
__global__ __launch_bounds__(64) void stddev_float_rows_kernel(const float* __restrict__ gray, long gp,
                                                               const uint8_t* __restrict__ alpha, long ap,
                                                               int w, int h, int n, float* __restrict__ out, long op)
{
    const int b = n / 2;
    const int row = blockIdx.x * 64 + threadIdx.x;
    if (row >= h - 2 * b) return;
    const int yc = row + b;
    float cs[NMAX], cq[NMAX];
    int cn[NMAX];
    auto column = [&](int x, float& s, float& q, int& k) {
        s = 0.f; q = 0.f; k = 0;
        for (int y = yc - b; y <= yc + b; ++y)
            if (alpha[(long)y * ap + x]) {
                const float v = gray[(long)y * gp + x];
                s += v;
                q += v * v;
                ++k;
            }
    };
    float sum = 0.f, sq = 0.f;
    int cnt = 0;
    for (int x = 0; x < n; ++x) {
        column(x, cs[x], cq[x], cn[x]);
        sum += cs[x];
        sq += cq[x];
        cnt += cn[x];
    }
    int slot = 0;  // ring position of the oldest column
    for (int xc = b;; ++xc) {
        if (alpha[(long)yc * ap + xc]) {
            float r = 0.f;
            if (cnt > 1) r = (float)sqrt((double)((sq - sum * sum / (float)cnt) / (float)(cnt - 1)));
            out[(long)yc * op + xc] = r;
        }
        if (xc == w - b - 1) break;
        float s, q;
        int k;
        column(xc + b + 1, s, q, k);
        sum += s - cs[slot];
        sq += q - cq[slot];
        cnt += k - cn[slot];
        cs[slot] = s; cq[slot] = q; cn[slot] = k;
        slot = slot + 1 == n ? 0 : slot + 1;
    }
}

// localStdDevIf for float gray images over LDS column chunks, several
// images per launch (blockIdx.y).  A block of four waves owns 64 consecutive
// window rows and walks the image in chunks of 64 columns:
//  * load: the chunk's samples, with the window's rows above and below, go
//    to LDS as (value under alpha or +0, 1 or 0) pairs (the next chunk is
//    already in flight in registers while the current one is processed);
//  * column sums (all waves, lane = column): each window row's column sum of
//    values and of squares, top to bottom, and its sample count, with no
//    branch: a sample outside alpha adds +0, which leaves every sum unchanged
//    bit for bit (a sum that starts at +0 never becomes -0);
//  * running sums (one wave, lane = row): the float running sums advance
//    column by column exactly in the reference's order, sum += s - (the
//    column leaving the window), the leaving column being zero while the
//    window fills; the column sums of the chunk sit N slots to the right of
//    the running sums, so step i reads slots N + i (entering column) and i
//    (leaving column, or the previous chunk's last N columns, kept in
//    registers, for i < N) and writes its running sums to slot i, which no
//    later step reads; the steps issue their loads ahead of the chain;
//  * results (all waves, lane = column): the standard deviation of every
//    window whose centre lies under alpha, stored as coalesced row segments.
// This is the only serial part of the path: one row is a chain of one float
// addition per column, so the other phases are spread over all four waves.
constexpr int kSdCW = 64;
constexpr int kSdMaxJobs = 8;

struct StddevJobs {
    const float* gray[kSdMaxJobs];
    const uint8_t* alpha[kSdMaxJobs];
    long apitch[kSdMaxJobs];
    float* out[kSdMaxJobs];
    long gpitch, opitch;
};

template <int N>
__global__ __launch_bounds__(256) void stddev_float_tile_kernel(StddevJobs J, int w, int h)
{
    constexpr int B = N / 2, TH = 64 + 2 * B, RPW = (TH + 3) / 4, CS = N + kSdCW;
    __shared__ float2 sg[TH][kSdCW + 1];             // (value under alpha or +0, alpha flag)
    __shared__ float css[64][CS + 1], csq[64][CS + 1];
    __shared__ uint8_t csk[64][CS + 3];
    __shared__ uint8_t prev_a[2][64][4];
    const int job = blockIdx.y;
    const float* __restrict__ gray = J.gray[job];
    const uint8_t* __restrict__ alpha = J.alpha[job];
    float* __restrict__ out = J.out[job];
    const long gp = J.gpitch, ap = J.apitch[job], op = J.opitch;
    // the wave index is uniform: row addresses stay scalar
    const int tid = threadIdx.x, wv = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int row0 = blockIdx.x * 64;
    float pg[RPW];
    uint8_t pa[RPW];
    auto fetch = [&](int xb) {
        const int gx = xb + lane;
#pragma unroll
        for (int u = 0; u < RPW; ++u) {
            const int r = wv + 4 * u, gy = row0 + r;
            const bool in = r < TH && gy < h && gx < w;
            pg[u] = in ? gray[(long)gy * gp + gx] : 0.f;
            pa[u] = in ? alpha[(long)gy * ap + gx] : (uint8_t)0;
        }
    };
    // chain state (wave 0): running sums and the previous chunk's last N
    // column sums
    float sum = 0.f, sq = 0.f;
    int cnt = 0;
    float ps[N], pq[N];
    int pk[N];
#pragma unroll
    for (int i = 0; i < N; ++i) { ps[i] = 0.f; pq[i] = 0.f; pk[i] = 0; }
    prev_a[0][lane][wv] = 0;
    fetch(0);
    int par = 0;
    for (int xb = 0; xb < w; xb += kSdCW, par ^= 1) {
#pragma unroll
        for (int u = 0; u < RPW; ++u) {
            const int r = wv + 4 * u;
            if (r < TH) sg[r][lane] = make_float2(pa[u] ? pg[u] : 0.f, pa[u] ? 1.f : 0.f);
        }
        __syncthreads();
        if (xb + kSdCW < w) fetch(xb + kSdCW);
        const int nx = min(kSdCW, w - xb);
        // column sums of window rows wv*16 .. wv*16+15, column `lane`
#pragma unroll 4
        for (int rr = 0; rr < 16; ++rr) {
            const int r = wv * 16 + rr;
            float s = 0.f, q = 0.f, k = 0.f;
#pragma unroll
            for (int dy = 0; dy < N; ++dy) {
                const float2 v = sg[r + dy][lane];
                s += v.x;
                q += v.x * v.x;
                k += v.y;
            }
            css[r][N + lane] = s;
            csq[r][N + lane] = q;
            csk[r][N + lane] = (uint8_t)k;
        }
        __syncthreads();
        if (wv == 0) {
#pragma unroll
            for (int i = 0; i < N; ++i) {
                if (i < nx) {
                    const float s = css[lane][N + i], q = csq[lane][N + i];
                    const int k = csk[lane][N + i];
                    sum += s - ps[i];
                    sq += q - pq[i];
                    cnt += k - pk[i];
                    css[lane][i] = sum;
                    csq[lane][i] = sq;
                    csk[lane][i] = (uint8_t)cnt;
                }
            }
#pragma unroll 8
            for (int i = N; i < nx; ++i) {
                const float s = css[lane][N + i], q = csq[lane][N + i];
                const int k = csk[lane][N + i];
                const float so = css[lane][i], qo = csq[lane][i];
                const int ko = csk[lane][i];
                sum += s - so;
                sq += q - qo;
                cnt += k - ko;
                css[lane][i] = sum;
                csq[lane][i] = sq;
                csk[lane][i] = (uint8_t)cnt;
            }
            if (nx == kSdCW) {
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    ps[j] = css[lane][kSdCW + j];
                    pq[j] = csq[lane][kSdCW + j];
                    pk[j] = csk[lane][kSdCW + j];
                }
            }
        }
        __syncthreads();
        // Column i = xb + lane completes the window centred on xb + lane - B;
        // for lane < B that centre lies in the previous chunk, whose last B
        // columns of centre alpha were kept in prev_a.
        for (int rr = 0; rr < 16; ++rr) {
            const int r = wv * 16 + rr;
            const int y = row0 + r + B, x = xb + lane - B;
            if (lane < nx && y < h - B && x >= B && x < w - B) {
                const bool ca = lane >= B ? sg[r + B][lane - B].y != 0.f : prev_a[par][r][lane] != 0;
                float res = 0.f;   // pixels outside the alpha keep localStdDevIf's zero
                if (ca) {
                    const int c = csk[r][lane];
                    const float sm = css[r][lane], sqv = csq[r][lane];
                    if (c > 1) res = (float)sqrt((double)((sqv - sm * sm / (float)c) / (float)(c - 1)));
                }
                out[(long)y * op + x] = res;
            }
        }
        for (int rr = 0; rr < 16; ++rr) {
            const int r = wv * 16 + rr;
            if (lane < B) prev_a[par ^ 1][r][lane] = sg[r + B][kSdCW - B + lane].y != 0.f ? 1 : 0;
        }
        __syncthreads();
    }
}

template <class T> __device__ __forceinline__ float contrast_term(double wc, double g);
template <> __device__ __forceinline__ float contrast_term<uint8_t>(double wc, double g) { return (float)(wc * g / 255.0); }
template <> __device__ __forceinline__ float contrast_term<uint16_t>(double wc, double g) { return (float)(wc * g / 65535.0); }
template <> __device__ __forceinline__ float contrast_term<float>(double wc, double g) { return (float)(wc * g); }

template <class T> struct GradT { typedef int32_t type; };
template <> struct GradT<float> { typedef float type; };

// enfuseMask's weight of one pixel under alpha (src/enfuse.h:1099-1211):
// exposure, then + contrast, + saturation, + entropy, float additions in the
// reference's order.
template <class T>
__device__ __forceinline__ float pixel_weight(const T (&p)[3], const WeightArgs& a, const float* __restrict__ exp_lut,
                                              typename GradT<T>::type gradv, T entv)
{
    float m = 0.f;
    // CutoffExposureFunctor with the default projectors: a pixel whose
    // minimum channel is below the lower cutoff or whose maximum channel
    // is above the upper one weighs 0
    const bool cut_out = a.exposure_cutoff &&
                         !((double)min(p[0], min(p[1], p[2])) >= a.exposure_lo &&
                           (double)max(p[0], max(p[1], p[2])) <= a.exposure_hi);
    if (a.w_exposure > 0.0 && !cut_out) {
        if (Img<T>::integral) {
            m = exp_lut[(int)Gray<T, T>::get(p, a.proj)];
        } else {
            const float g = Gray<T, float>::get(p, a.proj);
            const double yv = (double)(g / FLT_MAX);
            m = (float)(a.w_exposure * exposure_builtin(a.exposure_function, a.exposure_optimum, a.exposure_width, yv));
        }
    }
    if (a.w_contrast > 0.0) m = contrast_term<T>(a.w_contrast, (double)gradv) + m;
    if (a.w_saturation > 0.0) {
        const T mx = max(p[0], max(p[1], p[2]));
        const T mn = min(p[0], min(p[1], p[2]));
        float s = 0.f;
        if (!(mx == mn)) {
            const double max_value = Img<T>::integral ? Img<T>::maxv() : 1.0;
            const double sum = (double)mx + (double)mn;
            const double diff = (double)mx - (double)mn;
            s = (float)(a.w_saturation * (sum <= max_value ? diff / sum : diff / (2.0 * max_value - sum)));
        }
        m = s + m;
    }
    if (a.w_entropy > 0.0) m = (float)(a.w_entropy * (double)(float)entv) + m;
    return m;
}

// Outputs of the weights pass that the entropy kernel takes over when it
// computes the whole weight of the pixels it already holds.
template <class T> struct FusedWeights {
    WeightArgs a;
    const float* exp_lut;
    const typename GradT<T>::type* grad; long gp;
    float* mask; long mp;
    float* norm; long np_;
    uint8_t* ualpha; long up;
};

// The border of a float standard-deviation plane that localStdDevIf leaves
// at zero (rows and columns within n/2 of the edge): the tile kernel writes
// every interior pixel itself.  One block per row of one plane.
__global__ __launch_bounds__(64) void stddev_border_kernel(StddevJobs J, int w, int h, int b)
{
    const int y = blockIdx.x, job = blockIdx.y;
    float* row = J.out[job] + (long)y * J.opitch;
    const bool whole = y < b || y >= h - b;
    if (whole) {
        for (int x = threadIdx.x; x < w; x += 64) row[x] = 0.f;
    } else if ((int)threadIdx.x < 2 * b) {
        const int x = (int)threadIdx.x < b ? (int)threadIdx.x : w - 2 * b + (int)threadIdx.x;
        if (x >= 0 && x < w) row[x] = 0.f;
    }
}

// localEntropyIf + EntropyFunctor's channel minimum (src/enfuse.h:222-566,
// :832-861): minimum over channels of fromRealPromote(H_c * max(T)).
template <class T, int NMAX2>
__global__ __launch_bounds__(kTile) void entropy_kernel(const T* __restrict__ img, long ip,
                                                        const uint8_t* __restrict__ alpha, long ap,
                                                        int w, int h, int n, const double* __restrict__ tab,
                                                        T* __restrict__ out, long op)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const int b = n / 2;
    T res = T(0);
    if (x >= b && x < w - b && y >= b && y < h - b && alpha[(long)y * ap + x]) {
        const int full = n * n;
        // tab layout: [0, full]            p*log(p) for a full window (pent)
        //             [full+1, 2full+1]     log(bins) (plog)
        //             [2full+2, ...)        p*log(p) for total t, count k at (t * (full + 1) + k)
        const double* pent = tab;
        const double* plog = tab + (full + 1);
        const double* pgen = tab + 2 * (full + 1);
        T emin = T(0);
        for (int c = 0; c < 3; ++c) {
            T v[NMAX2];
            int cnt = 0;
            for (int yy = y - b; yy <= y + b; ++yy)
                for (int xx = x - b; xx <= x + b; ++xx)
                    if (alpha[(long)yy * ap + xx]) {
                        const T s = img[(long)yy * ip + 3L * xx + c];
                        int k = cnt++;
                        while (k > 0 && s < v[k - 1]) { v[k] = v[k - 1]; --k; }
                        v[k] = s;
                    }
            double e = 0.0;
            if (cnt > 0) {
                int bins = 0;
                double acc = 0.0;
                int run = 1;
                for (int k = 1; k <= cnt; ++k) {
                    if (k < cnt && !(v[k - 1] < v[k])) { ++run; continue; }
                    acc += cnt == full ? pent[run] : pgen[cnt * (full + 1) + run];
                    ++bins;
                    run = 1;
                }
                if (bins > 1) e = -acc / plog[bins];
            }
            const T ec = from_real_img<T>(e * Img<T>::maxv());
            emin = c == 0 ? ec : (ec < emin ? ec : emin);
        }
        res = emin;
    }
    out[(long)y * op + x] = res;
}

// ---- local entropy over LDS tiles ------------------------------------------
//
// A block of 256 threads owns a 64 x 8 tile of output pixels (two rows per
// thread).  The tile's samples with an n/2 halo (channel-planar) and their
// alpha go to LDS once, then the minimum and maximum of every n-sample window
// column.  Per pixel and channel, from those column extremes alone:
//   * a window of one value (minimum == maximum) is one bin: entropy 0, so
//     the pixel's channel minimum is 0 and nothing else is computed;
//   * a window whose n columns are each constant has one bin per distinct
//     column value, of n samples per column: its n column keys are sorted
//     and walked like the full window's keys would be;
//   * any other window goes to a per-block queue.  After a barrier every
//     thread takes queue entries in turn, sorts the entry's n*n keys by a
//     comparator network held in registers (Batcher's merge exchange) and
//     walks the bins of equal keys in ascending key order,
//     accumulating the full-window p*log(p) table in the order the
//     reference's std::map iterates them.  So the sorting work of a block is
//     spread over all its lanes instead of being paid by a whole wave for
//     any one lane that needs it.
// A tile whose halo lies wholly under alpha and inside the image and has a
// channel with one value throughout (every sample equal to the tile's first
// one) has a one-bin window in that channel at every pixel, so its entropies
// are all 0: such a tile skips the column extremes, classes and queues.
// Windows that are cut by alpha take an insertion sort of the samples under
// alpha and the general table, in place.  Equal keys are those that compare
// neither less nor greater (the map's equivalence), so -0 and +0 share a bin.
// NaN samples break the map's strict weak ordering in the reference, so their
// bins are not defined there.

struct CmpNet {
    int n;
    unsigned char a[1024], b[1024];
};

// Batcher's merge exchange for N keys (Knuth, TAOCP vol. 3, 5.2.2,
// Algorithm M): a sorting network for any N, with no comparator past the
// last key (383 comparators for 49 keys, against 394 for the odd-even merge
// sort of 64 pruned to 49).
constexpr CmpNet make_net(int N)
{
    CmpNet c{};
    int t = 1;
    while ((1 << t) < N) ++t;
    int m = 0;
    for (int p = 1 << (t - 1); p > 0; p >>= 1) {
        int q = 1 << (t - 1), r = 0, d = p;
        for (;;) {
            for (int i = 0; i < N - d; ++i)
                if ((i & p) == r) {
                    c.a[m] = (unsigned char)i;
                    c.b[m] = (unsigned char)(i + d);
                    ++m;
                }
            if (q == p) break;
            d = q - p;
            q >>= 1;
            r = p;
        }
    }
    c.n = m;
    return c;
}

template <int NN> struct SortNet { static constexpr CmpNet net = make_net(NN); };

template <class T> __device__ __forceinline__ void cas(T& x, T& y) { const T lo = min(x, y), hi = max(x, y); x = lo; y = hi; }
template <class T> __device__ __forceinline__ void kmin(T& m, T v) { m = min(m, v); }
template <class T> __device__ __forceinline__ void kmax(T& m, T v) { m = max(m, v); }
template <> __device__ __forceinline__ void kmin<float>(float& m, float v) { m = __builtin_fminf(m, v); }
template <> __device__ __forceinline__ void kmax<float>(float& m, float v) { m = __builtin_fmaxf(m, v); }
template <> __device__ __forceinline__ void cas<float>(float& x, float& y)
{
    const float lo = __builtin_fminf(x, y), hi = __builtin_fmaxf(x, y);
    x = lo;
    y = hi;
}

// Integer samples are sorted as int (u8/u16 values), float samples as float.
template <class T> struct SortKey { typedef int type; };
template <> struct SortKey<float> { typedef float type; };

constexpr int kEntTW = 64, kEntTH = 8, kEntThreads = 256;

// Bins of the ascending keys k[0..M-1] (at least two distinct), each key
// standing for `mult` samples: adds the full-window table term of every bin
// in ascending key order.  Non-boundaries add spent[0] = 0.0, which leaves the
// sum unchanged bit for bit, so the table reads do not wait on one another.
template <class K, int M>
__device__ __forceinline__ double walk_bins(const K (&k)[M], int mult, const double* spent, int& bins)
{
    double acc = 0.0;
    int st = 0;
    bins = 1;
#pragma unroll
    for (int i = 1; i < M; ++i) {
        const bool b = k[i - 1] < k[i];
        const int len = i - st;
        st = b ? i : st;
        bins += b ? 1 : 0;
        acc += spent[b ? len * mult : 0];
    }
    return acc + spent[(M - st) * mult];
}

// Appends the lanes with `pred` set to a block queue with one LDS atomic per
// wave; returns each such lane's slot.  Called by every lane of the wave.
__device__ __forceinline__ int wave_append(int* counter, bool pred)
{
    const unsigned long long m = __ballot(pred);
    if (m == 0ull) return -1;
    const int lane = (int)(threadIdx.x & 63u);
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader);
    return pred ? base + rank : -1;
}

// FUSEW: the kernel also makes the pixel's whole weight (the weights pass
// of pixel_weight, with the entropy it has just computed and the pixel and
// alpha it holds in LDS) instead of writing the entropy plane.
template <class T, int N, bool FUSEW>
__global__ __launch_bounds__(kEntThreads) void entropy_tile_kernel(const T* __restrict__ img, long ip,
                                                                   const uint8_t* __restrict__ alpha, long ap,
                                                                   int w, int h, const double* __restrict__ tab,
                                                                   T* __restrict__ out, long op, FusedWeights<T> fw)
{
    constexpr int B = N / 2, NN = N * N;
    constexpr int LW = kEntTW + 2 * B, LH = kEntTH + 2 * B;
    constexpr int NT = kEntThreads, NW = NT / 64, RPT = kEntTW * kEntTH / NT, RSTEP = NT / kEntTW;
    constexpr int WROWS = (LH + NW - 1) / NW, VCH = (3 * LW + 63) / 64, ACH = (LW + 63) / 64;
    constexpr int QMAX = 3 * kEntTW * kEntTH;
    typedef typename SortKey<T>::type K;
    __shared__ K sv[3][LH][LW];
    __shared__ K cmn[3][kEntTH][LW];
    // the column maxima are dead once the window classes are made; the
    // channel entropies take their place
    __shared__ union {
        K cmx[3][kEntTH][LW];
        T ent[3][kEntTH][kEntTW];
    } u;
    __shared__ uint8_t cls[3][kEntTH][kEntTW];
    __shared__ uint8_t cmo[3][kEntTH][LW];   // window column strictly monotone
    __shared__ uint8_t sa[LH][LW + 4];
    __shared__ double spent[NN + 1], splog[NN + 1];
    __shared__ unsigned short queue[QMAX], cqueue[QMAX];
    __shared__ int qn, cqn;
    __shared__ int wbits[NW];
    auto& cmx = u.cmx;
    auto& ent = u.ent;
    const int x0 = blockIdx.x * kEntTW, y0 = blockIdx.y * kEntTH;
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    // Halo tile, one wave per tile row (rows wv, wv + NW, ...), lanes along
    // the row's interleaved samples; samples outside the image count as
    // alpha 0.  All loads of a thread are issued before the first LDS store.
    K lv[WROWS][VCH];
    uint8_t la[WROWS][ACH];
    int voff[VCH];
    bool vin[VCH];
#pragma unroll
    for (int j = 0; j < VCH; ++j) {
        const int e = lane + 64 * j;
        const int lx = e / 3, c = e - 3 * lx;
        const int gx = x0 - B + lx;
        vin[j] = e < 3 * LW && gx >= 0 && gx < w;
        voff[j] = c * (LH * LW) + lx;
    }
#pragma unroll
    for (int k = 0; k < WROWS; ++k) {
        const int ly = wv + NW * k, gy = y0 - B + ly;
        const bool rin = ly < LH && gy >= 0 && gy < h;
        const T* row = img + (long)gy * ip + 3L * (x0 - B);
        const uint8_t* arow = alpha + (long)gy * ap + (x0 - B);
#pragma unroll
        for (int j = 0; j < VCH; ++j) lv[k][j] = rin && vin[j] ? (K)row[lane + 64 * j] : (K)0;
#pragma unroll
        for (int j = 0; j < ACH; ++j) {
            const int lx = lane + 64 * j, gx = x0 - B + lx;
            la[k][j] = rin && lx < LW && gx >= 0 && gx < w ? arow[lx] : (uint8_t)0;
        }
    }
    // the tile's first halo sample per channel (NaN-free keys compare equal
    // to themselves; a NaN reference leaves the channel unflagged)
    K ref[3];
    const bool corner = x0 - B >= 0 && y0 - B >= 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) ref[c] = corner ? (K)img[(long)(y0 - B) * ip + 3L * (x0 - B) + c] : (K)0;
    int unset = corner ? 0 : 1;
    int diff = 0;   // bit c: a sample of channel c differs from ref[c]
#pragma unroll
    for (int k = 0; k < WROWS; ++k) {
        const int ly = wv + NW * k;
        if (ly >= LH) continue;
#pragma unroll
        for (int j = 0; j < VCH; ++j)
            if (lane + 64 * j < 3 * LW) {
                (&sv[0][0][0])[voff[j] + ly * LW] = lv[k][j];
                const int c = voff[j] / (LH * LW);
                const K rv = c == 0 ? ref[0] : (c == 1 ? ref[1] : ref[2]);
                diff |= lv[k][j] == rv ? 0 : (1 << c);
            }
#pragma unroll
        for (int j = 0; j < ACH; ++j) {
            const int lx = lane + 64 * j;
            if (lx < LW) {
                sa[ly][lx] = la[k][j];
                unset |= la[k][j] == 0;
            }
        }
    }
    for (int i = tid; i <= NN; i += NT) {
        spent[i] = tab[i];
        splog[i] = tab[NN + 1 + i];
    }
    if (tid == 0) { qn = 0; cqn = 0; }
    const double sdist = tab[2 * (NN + 1) + (NN + 1) * (NN + 1)];
    {
        const int bits = (unset ? 8 : 0) | diff;
        int wb = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) wb |= __ballot((bits >> q) & 1) ? (1 << q) : 0;
        if (lane == 0) wbits[wv] = wb;
    }
    __syncthreads();
    int tbits = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) tbits |= wbits[q];
    const bool tile_full = (tbits & 8) == 0;
    // a channel with one value over the whole tile: every entropy is 0
    const bool flat_tile = tile_full && (tbits & 7) != 7;
    const int lx = tid % kEntTW;
    T res[RPT];
    int pend[RPT];                 // bit c: channel c's value comes from a queue
    bool store[RPT];
#pragma unroll
    for (int p = 0; p < RPT; ++p) {
        res[p] = T(0);
        pend[p] = 0;
        store[p] = x0 + lx < w && y0 + tid / kEntTW + p * RSTEP < h;
    }
    if (!flat_tile) {
        // Extremes of every window column (N rows from output row r): one thread
        // per tile column and channel reads the column once.
        for (int i = tid; i < 3 * LW; i += NT) {
            const int c = i / LW, cx = i - c * LW;
            K col[LH];
    #pragma unroll
            for (int r = 0; r < LH; ++r) col[r] = sv[c][r][cx];
            bool lt[LH - 1], gt[LH - 1];
    #pragma unroll
            for (int r = 0; r + 1 < LH; ++r) {
                lt[r] = col[r] < col[r + 1];
                gt[r] = col[r + 1] < col[r];
            }
    #pragma unroll
            for (int r = 0; r < kEntTH; ++r) {
                K mn = col[r], mx = col[r];
                bool inc = true, dec = true;
    #pragma unroll
                for (int dy = 1; dy < N; ++dy) {
                    kmin(mn, col[r + dy]);
                    kmax(mx, col[r + dy]);
                    inc = inc && lt[r + dy - 1];
                    dec = dec && gt[r + dy - 1];
                }
                cmn[c][r][cx] = mn;
                cmx[c][r][cx] = mx;
                cmo[c][r][cx] = (uint8_t)(inc || dec);
            }
        }
        __syncthreads();
        // Window classes from the column extremes, sliding along each row in
        // runs of kEntSeg outputs: bit 0 = one value (flat), bit 1 = every
        // column constant, bit 2 = every column strictly monotone and the
        // column ranges strictly ordered along the row (so the N*N samples
        // are distinct: N*N bins of one sample, no sort needed).
        constexpr int kEntSeg = 8, NSEG = kEntTW / kEntSeg;
        for (int i = tid; i < 3 * kEntTH * NSEG; i += NT) {
            const int c = i / (kEntTH * NSEG), e = i - c * (kEntTH * NSEG);
            const int r = e / NSEG, x0s = (e - r * NSEG) * kEntSeg;
            K rmn[kEntSeg + N - 1], rmx[kEntSeg + N - 1];
            bool rmo[kEntSeg + N - 1], up[kEntSeg + N - 2], dn[kEntSeg + N - 2];
    #pragma unroll
            for (int j = 0; j < kEntSeg + N - 1; ++j) {
                rmn[j] = cmn[c][r][x0s + j];
                rmx[j] = cmx[c][r][x0s + j];
                rmo[j] = cmo[c][r][x0s + j] != 0;
            }
    #pragma unroll
            for (int j = 0; j + 1 < kEntSeg + N - 1; ++j) {
                up[j] = rmx[j] < rmn[j + 1];
                dn[j] = rmx[j + 1] < rmn[j];
            }
    #pragma unroll
            for (int j = 0; j < kEntSeg; ++j) {
                K mn = rmn[j], mx = rmx[j];
                bool cols = !(rmn[j] < rmx[j]);
                bool mono = rmo[j], ou = true, od = true;
    #pragma unroll
                for (int dx = 1; dx < N; ++dx) {
                    kmin(mn, rmn[j + dx]);
                    kmax(mx, rmx[j + dx]);
                    cols = cols && !(rmn[j + dx] < rmx[j + dx]);
                    mono = mono && rmo[j + dx];
                    ou = ou && up[j + dx - 1];
                    od = od && dn[j + dx - 1];
                }
                const bool ordered = mono && (ou || od);
                cls[c][r][x0s + j] = (uint8_t)((mn < mx ? 0 : 1) | (cols ? 2 : 0) | (ordered ? 4 : 0));
            }
        }
        __syncthreads();
        // Per pixel: cut windows in place; a flat channel makes the minimum 0;
        // every other channel goes to the queue of constant-column windows or
        // to the sorting queue.
    #pragma unroll
        for (int p = 0; p < RPT; ++p) {
            const int ly = tid / kEntTW + p * RSTEP;
            const int x = x0 + lx, y = y0 + ly;
            res[p] = T(0);
            pend[p] = 0;
            store[p] = x < w && y < h;
            if (!(store[p] && x >= B && x < w - B && y >= B && y < h - B && sa[ly + B][lx + B])) continue;
            bool full = tile_full;
            if (!full) {
                full = true;
    #pragma unroll
                for (int dy = 0; dy < N; ++dy)
    #pragma unroll
                    for (int dx = 0; dx < N; ++dx) full = full && sa[ly + dy][lx + dx] != 0;
            }
            if (!full) {
                // Cut window: insertion sort of the samples under alpha; the
                // general p*log(p) table for `cnt` samples.
                const double* pgen = tab + 2 * (NN + 1);
                T emin = T(0);
                for (int c = 0; c < 3; ++c) {
                    K v[NN];
                    int cnt = 0;
                    for (int dy = 0; dy < N; ++dy)
                        for (int dx = 0; dx < N; ++dx)
                            if (sa[ly + dy][lx + dx]) {
                                const K sm = sv[c][ly + dy][lx + dx];
                                int k = cnt++;
                                while (k > 0 && sm < v[k - 1]) { v[k] = v[k - 1]; --k; }
                                v[k] = sm;
                            }
                    double acc = 0.0;
                    int bins = 0, run = 1;
                    for (int k = 1; k <= cnt; ++k) {
                        if (k < cnt && !(v[k - 1] < v[k])) { ++run; continue; }
                        acc += pgen[cnt * (NN + 1) + run];
                        ++bins;
                        run = 1;
                    }
                    const double e = bins > 1 ? -acc / splog[bins] : 0.0;
                    const T ec = from_real_img<T>(e * Img<T>::maxv());
                    emin = c == 0 ? ec : (ec < emin ? ec : emin);
                    // every channel's value is +0 or more: once the minimum is 0
                    // the remaining channels cannot change it
                    if (emin == T(0)) break;
                }
                res[p] = emin;
                continue;
            }
            const int k0 = cls[0][ly][lx], k1 = cls[1][ly][lx], k2 = cls[2][ly][lx];
            if ((k0 | k1 | k2) & 1) continue;     // a one-bin channel: the minimum is 0
            pend[p] = 7 | (((k0 >> 1) & 1) << 3) | (((k1 >> 1) & 1) << 4) | (((k2 >> 1) & 1) << 5) |
                      (((k0 >> 2) & 1) << 6) | (((k1 >> 2) & 1) << 7) | (((k2 >> 2) & 1) << 8);
        }
        // the entropy of N*N distinct samples (the sorting queue's distinct
        // windows take it too)
        const T ent_dist = from_real_img<T>(-sdist / splog[NN] * Img<T>::maxv());
    #pragma unroll
        for (int p = 0; p < RPT; ++p) {
            const int pix = (tid / kEntTW + p * RSTEP) * kEntTW + lx;
    #pragma unroll
            for (int c = 0; c < 3; ++c) {
                const bool ordered = (pend[p] >> (6 + c)) & 1;
                const bool queued = ((pend[p] >> c) & 1) && !ordered;
                const bool colc = (pend[p] >> (3 + c)) & 1;
                if (ordered) ent[c][pix / kEntTW][lx] = ent_dist;
                const int cs = wave_append(&cqn, queued && colc);
                const int ss = wave_append(&qn, queued && !colc);
                if (cs >= 0) cqueue[cs] = (unsigned short)((pix << 2) | c);
                if (ss >= 0) queue[ss] = (unsigned short)((pix << 2) | c);
            }
        }
        __syncthreads();
        // Constant-column windows: one bin per distinct column value, N samples
        // per column; the N column keys are sorted and walked.
        const int ncq = cqn;
        for (int i = tid; i < ncq; i += NT) {
            const int item = cqueue[i];
            const int c = item & 3, pix = item >> 2;
            const int qy = pix / kEntTW, qx = pix - qy * kEntTW;
            K k[N];
    #pragma unroll
            for (int dx = 0; dx < N; ++dx) k[dx] = cmn[c][qy][qx + dx];
            constexpr CmpNet net = SortNet<N>::net;
    #pragma unroll
            for (int q = 0; q < net.n; ++q) cas(k[net.a[q]], k[net.b[q]]);
            int bins;
            const double acc = walk_bins(k, N, spent, bins);
            ent[c][qy][qx] = from_real_img<T>(-acc / splog[bins] * Img<T>::maxv());
        }
        // Other windows: all NN keys sorted.
        const int nq = qn;
        for (int i = tid; i < nq; i += NT) {
            const int item = queue[i];
            const int c = item & 3, pix = item >> 2;
            const int qy = pix / kEntTW, qx = pix - qy * kEntTW;
            K v[NN];
    #pragma unroll
            for (int dy = 0; dy < N; ++dy)
    #pragma unroll
                for (int dx = 0; dx < N; ++dx) v[dy * N + dx] = sv[c][qy + dy][qx + dx];
            constexpr CmpNet net = SortNet<NN>::net;
    #pragma unroll
            for (int q = 0; q < net.n; ++q) cas(v[net.a[q]], v[net.b[q]]);
            bool distinct = true;
    #pragma unroll
            for (int k = 1; k < NN; ++k) distinct = distinct && v[k - 1] < v[k];
            // NN bins of one sample: the table's one-sample term added NN
            // times in order, precomputed with its entropy; the walk over the
            // bins runs only in waves that hold a window with tied samples
            T ev = ent_dist;
            if (__ballot(!distinct) != 0ull) {
                int bins;
                const double acc = walk_bins(v, 1, spent, bins);
                const double e = bins > 1 ? -acc / splog[bins] : 0.0;
                if (!distinct) ev = from_real_img<T>(e * Img<T>::maxv());
            }
            ent[c][qy][qx] = ev;
        }
        __syncthreads();
    }
    // Channel minimum (entropies are +0 or more and never NaN, so the order
    // of the comparisons does not matter).
#pragma unroll
    for (int p = 0; p < RPT; ++p) {
        if (!store[p]) continue;
        const int ly = tid / kEntTW + p * RSTEP;
        T r = res[p];
        if (pend[p] & 7) {
            r = ent[0][ly][lx];
            const T e1 = ent[1][ly][lx], e2 = ent[2][ly][lx];
            r = e1 < r ? e1 : r;
            r = e2 < r ? e2 : r;
        }
        const long y = y0 + ly, x = x0 + lx;
        if constexpr (FUSEW) {
            const uint8_t al = sa[ly + B][lx + B];
            float m = 0.f;
            if (al) {
                const T px[3] = {(T)sv[0][ly + B][lx + B], (T)sv[1][ly + B][lx + B], (T)sv[2][ly + B][lx + B]};
                m = pixel_weight<T>(px, fw.a, fw.exp_lut, fw.a.w_contrast > 0.0 ? fw.grad[y * fw.gp + x] : 0, r);
                fw.ualpha[y * fw.up + x] = al;
            }
            fw.mask[y * fw.mp + x] = m;
            float* nrm = fw.norm + y * fw.np_ + x;
            *nrm = m + *nrm;
        } else {
            out[y * op + x] = r;
        }
    }
}

// Combines the weights of one image in the reference's order (exposure,
// contrast, saturation, entropy; float additions), and updates the norm
// image and the union alpha (src/enfuse.h:1377-1385).
template <class T>
__global__ __launch_bounds__(kTile) void weights_kernel(const T* __restrict__ img, long ip,
                                                        const uint8_t* __restrict__ alpha, long ap, int w,
                                                        WeightArgs a,
                                                        const float* __restrict__ exp_lut,
                                                        const typename GradT<T>::type* __restrict__ grad, long gp,
                                                        const T* __restrict__ ent, long ep,
                                                        float* __restrict__ mask, long mp,
                                                        float* __restrict__ norm, long np_,
                                                        uint8_t* __restrict__ ualpha, long up)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const uint8_t al = alpha[(long)y * ap + x];
    float m = 0.f;
    if (al) {
        const T* p = img + (long)y * ip + 3L * x;
        const T px[3] = {p[0], p[1], p[2]};
        m = pixel_weight<T>(px, a, exp_lut, a.w_contrast > 0.0 ? grad[(long)y * gp + x] : 0,
                            a.w_entropy > 0.0 ? ent[(long)y * ep + x] : T(0));
        ualpha[(long)y * up + x] = al;
    }
    mask[(long)y * mp + x] = m;
    float* nrm = norm + (long)y * np_ + x;
    *nrm = m + *nrm;
}

// Normalisation (src/enfuse.h:1529-1534) fused with the fixed-point
// conversion of mask level 0 (src/fixmath.h:188-194); hard masks are
// converted as they are (the reference skips the normalisation).
template <class M>
__global__ __launch_bounds__(kTile) void normalize_kernel(const float* __restrict__ mask, long mp,
                                                          const float* __restrict__ norm, long np_, int w,
                                                          float fallback, double scale, int hard,
                                                          M* __restrict__ out, long op)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const float mv = mask[(long)y * mp + x];
    const float nv = norm[(long)y * np_ + x];
    const float v = hard ? mv : nv > 0.0 ? (255.0f * mv) / nv : fallback;
    out[(long)y * op + x] = from_real<M>((double)v * scale);
}

// Hard masks (src/enfuse.h:1402-1443) over the K weight planes: per pixel
// the first image whose weight is the largest above 0 gets 255, the others
// 0; where no weight is above 0, every image gets float(255) / K.
__global__ __launch_bounds__(kTile) void hard_mask_kernel(float* __restrict__ masks, long mp, size_t plane,
                                                          int w, int K)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    float* p = masks + (long)y * mp + x;
    float mx = 0.0f;
    int maxi = 0;
    for (int m = 0; m < K; ++m) {
        const float wv = p[m * plane];
        if (wv > mx) { mx = wv; maxi = m; }
    }
    const float fb = (float)255 / K;
    for (int m = 0; m < K; ++m) p[m * plane] = mx == 0.0f ? fb : (m == maxi ? 255.0f : 0.0f);
}

// ImageMaskMultiplyFunctor (src/enfuse.h:601-620) fused with the
// accumulation into the result pyramid (src/enfuse.h:1576-1588), for up to
// kMergeMax images at one level in one pass: per image the coefficient mask /
// max, the product rounded to the pyramid type, added with its wrap, in image
// order (the first of the call assigns when `first`), so the result keeps
// every rounding of the image-by-image accumulation while it is read and
// written once.
template <class P, class M>
__global__ __launch_bounds__(kTile) void merge_many_kernel(MergeMany a)
{
    typedef typename Pix<P>::S S;
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= a.w) return;
    P r[3];
    P* out = reinterpret_cast<P*>(a.res) + (long)y * a.rp + x;
    if (!a.first) {
#pragma unroll
        for (int c = 0; c < 3; ++c) r[c] = *gmem(out + c * a.rl);
    }
    for (int k = 0; k < a.n; ++k) {
        const double coeff = (double)*gmem(reinterpret_cast<const M*>(a.g[k]) + (long)y * a.gpp + x) / a.maxpyr;
        const P* l = reinterpret_cast<const P*>(a.lp[k]) + (long)y * a.lpp + x;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const P v = from_real<P>((double)*gmem(l + c * a.lpl) * coeff);
            r[c] = (a.first && k == 0) ? v : Pix<P>::wrap(Pix<P>::A::add((S)v, (S)r[c]));
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) *gmem(out + c * a.rl) = r[c];
}

// ---- contrast by Laplacian-of-Gaussian edges ------------------------------
// VIGRA's convolveLine with BORDER_TREATMENT_REFLECT: the window's samples
// from the leftmost (topmost) one, source index mirrored at both ends
// without repeating the edge sample, summed in double from 0 in that order;
// the symmetric kernels make tap t weigh k[t].  One thread per pixel; the
// taps of neighbouring threads overlap, so the re-reads come from cache.
__device__ __forceinline__ int mirror(int i, int n) { return i < 0 ? -i : (i >= n ? 2 * (n - 1) - i : i); }

template <class G> __device__ __forceinline__ G cast_real(double v);
template <> __device__ __forceinline__ int32_t cast_real<int32_t>(double v) { return from_real<int32_t>(v); }
template <> __device__ __forceinline__ float cast_real<float>(double v) { return (float)v; }
template <> __device__ __forceinline__ double cast_real<double>(double v) { return v; }

template <class G, class R, bool TWO>
__global__ __launch_bounds__(256) void edge_rows_kernel(EdgePlanes e, const double* __restrict__ ka, int ra,
                                                        const double* __restrict__ kb, int rb)
{
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= e.w) return;
    const G* row = reinterpret_cast<const G*>(e.src) + (long)y * e.spitch;
    double sa = 0.0;
    for (int t = 0; t <= 2 * ra; ++t) sa += ka[t] * (double)*gmem(row + mirror(x - ra + t, e.w));
    *gmem(reinterpret_cast<R*>(e.tA) + (long)y * e.tpitch + x) = cast_real<R>(sa);
    if (TWO) {
        double sb = 0.0;
        for (int t = 0; t <= 2 * rb; ++t) sb += kb[t] * (double)*gmem(row + mirror(x - rb + t, e.w));
        *gmem(reinterpret_cast<R*>(e.tB) + (long)y * e.tpitch + x) = cast_real<R>(sb);
    }
}

template <class G, class R>
__global__ __launch_bounds__(256) void edge_sharpen_kernel(EdgePlanes e, const double* __restrict__ k, int r, double f,
                                                           const G* __restrict__ gray, long gp, G* out, long op)
{
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= e.w) return;
    const R* col = reinterpret_cast<const R*>(e.tA) + x;
    double sm = 0.0;
    for (int t = 0; t <= 2 * r; ++t) sm += k[t] * (double)*gmem(col + (long)mirror(y - r + t, e.h) * e.tpitch);
    const R t2 = cast_real<R>(sm);
    const double v = (1.0 + f) * (double)*gmem(gray + (long)y * gp + x) - f * (double)t2;
    *gmem(out + (long)y * op + x) = cast_real<G>(v);
}

template <class G> __device__ __forceinline__ G g_max();
template <> __device__ __forceinline__ int32_t g_max<int32_t>() { return INT32_MAX; }
template <> __device__ __forceinline__ float g_max<float>() { return FLT_MAX; }

template <class G, class R, bool FILL>
__global__ __launch_bounds__(256) void edge_log_kernel(EdgePlanes e, const double* __restrict__ ks, int rs,
                                                       const double* __restrict__ kd, int rd,
                                                       const uint8_t* __restrict__ alpha, long ap,
                                                       const G* __restrict__ lc, long lp, double mc, double smax,
                                                       G* grad, long gp)
{
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= e.w) return;
    G out = (G)0;
    if (*gmem(alpha + (long)y * ap + x)) {
        const R* ca = reinterpret_cast<const R*>(e.tA) + x;   // rows: second derivative
        const R* cb = reinterpret_cast<const R*>(e.tB) + x;   // rows: smoothing
        double sx = 0.0, sy = 0.0;
        for (int t = 0; t <= 2 * rs; ++t) sx += ks[t] * (double)*gmem(ca + (long)mirror(y - rs + t, e.h) * e.tpitch);
        for (int t = 0; t <= 2 * rd; ++t) sy += kd[t] * (double)*gmem(cb + (long)mirror(y - rd + t, e.h) * e.tpitch);
        const R lap = (R)(cast_real<R>(sx) + cast_real<R>(sy));
        const G mag = cast_real<G>((double)fabs(lap));   // MagnitudeAccessor
        if (FILL) {
            // FillInFunctor(mc, 1, mc / max(ScalarType))
            const G thr = (G)mc;
            out = mag >= thr ? cast_real<G>(1.0 * (double)mag)
                             : cast_real<G>(mc / smax * (double)*gmem(lc + (long)y * lp + x));
        } else {
            // ClampingFunctor(-mc -> 0, max -> max)
            const G lo = (G)(-mc), up = g_max<G>();
            out = mag <= lo ? (G)0 : (mag >= up ? up : mag);
        }
    }
    *gmem(grad + (long)y * gp + x) = out;
}

// --entropy-cutoff: ClampingFunctor<RGB>(lo -> 0, up -> max) applied to the
// image before the local entropy; its blue channel tests the red channel
// against the lower bound, as the reference's does.
template <class T>
__global__ __launch_bounds__(kTile) void entropy_clamp_kernel(const T* __restrict__ img, long ip, int w, T lo, T up,
                                                              T mx, T* __restrict__ out)
{
    const int x = blockIdx.x * kTile + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const T* p = img + (long)y * ip + 3L * x;
    T* q = out + ((long)y * w + x) * 3;
    const T r = p[0], g = p[1], b = p[2];
    q[0] = r <= lo ? (T)0 : (r >= up ? mx : r);
    q[1] = g <= lo ? (T)0 : (g >= up ? mx : g);
    q[2] = r <= lo ? (T)0 : (b >= up ? mx : b);
}

}  // namespace

int launch_gray(int itype, const void* img, long ip, int w, int h, void* gray, long gp, const GrayProj& pj, hipStream_t s)
{
    if (itype == ENBL_U8) hipLaunchKernelGGL((gray_kernel<uint8_t, int32_t>), grid_for(w, h), dim3(kTile), 0, s, (const uint8_t*)img, ip, w, (int32_t*)gray, gp, pj);
    else if (itype == ENBL_U16) hipLaunchKernelGGL((gray_kernel<uint16_t, int32_t>), grid_for(w, h), dim3(kTile), 0, s, (const uint16_t*)img, ip, w, (int32_t*)gray, gp, pj);
    else if (itype == ENBL_F32) hipLaunchKernelGGL((gray_kernel<float, float>), grid_for(w, h), dim3(kTile), 0, s, (const float*)img, ip, w, (float*)gray, gp, pj);
    else return ENBL_E_UNSUPPORTED;
    return 0;
}

int launch_stddev(int itype, const void* gray, long gp, size_t gstride, const uint8_t* const* alphas,
                  const long* apitches, int K, int w, int h, int n, void* out, long op, size_t ostride,
                  hipStream_t s)
{
    if (itype == ENBL_F32) {
        if (n > 31) return ENBL_E_UNSUPPORTED;
        const int rows = h - 2 * (n / 2);
        if (n == 3 || n == 5 || n == 7) {
            for (int m0 = 0; m0 < K; m0 += kSdMaxJobs) {
                const int nj = min(kSdMaxJobs, K - m0);
                StddevJobs J;
                memset(&J, 0, sizeof J);
                J.gpitch = gp;
                J.opitch = op;
                for (int j = 0; j < nj; ++j) {
                    J.gray[j] = (const float*)gray + (m0 + j) * gstride;
                    J.alpha[j] = alphas[m0 + j];
                    J.apitch[j] = apitches[m0 + j];
                    J.out[j] = (float*)out + (m0 + j) * ostride;
                }
                if (rows > 0) {
                    const dim3 grid((rows + 63) / 64, nj), block(256);
                    if (n == 3) hipLaunchKernelGGL((stddev_float_tile_kernel<3>), grid, block, 0, s, J, w, h);
                    else if (n == 5) hipLaunchKernelGGL((stddev_float_tile_kernel<5>), grid, block, 0, s, J, w, h);
                    else hipLaunchKernelGGL((stddev_float_tile_kernel<7>), grid, block, 0, s, J, w, h);
                }
                hipLaunchKernelGGL(stddev_border_kernel, dim3(h, nj), dim3(64), 0, s, J, w, h, n / 2);
            }
            return 0;
        }
        if (rows <= 0) return 0;
        for (int m = 0; m < K; ++m)
            hipLaunchKernelGGL((stddev_float_rows_kernel<31>), dim3((rows + 63) / 64), dim3(64), 0, s,
                               (const float*)gray + m * gstride, gp, alphas[m], apitches[m], w, h, n,
                               (float*)out + m * ostride, op);
        return 0;
    }
    for (int m = 0; m < K; ++m)
        hipLaunchKernelGGL(stddev_int_kernel, grid_for(w, h), dim3(kTile), 0, s, (const int32_t*)gray + m * gstride, gp,
                           alphas[m], apitches[m], w, h, n, (int32_t*)out + m * ostride, op);
    return 0;
}

int launch_entropy(int itype, const void* img, long ip, const uint8_t* alpha, long ap, int w, int h, int n,
                   const double* tab, void* out, long op, hipStream_t s)
{
    if (n * n > 121) return ENBL_E_UNSUPPORTED;
    if (n == 3 || n == 5 || n == 7) {
        const dim3 grid((w + kEntTW - 1) / kEntTW, (h + kEntTH - 1) / kEntTH), block(kEntThreads);
#define ENBL_ENT_TILE(T, N) hipLaunchKernelGGL((entropy_tile_kernel<T, N, false>), grid, block, 0, s, (const T*)img, ip, alpha, ap, w, h, tab, (T*)out, op, FusedWeights<T>{})
#define ENBL_ENT_N(T)                   \
    do {                                \
        if (n == 3) ENBL_ENT_TILE(T, 3); \
        else if (n == 5) ENBL_ENT_TILE(T, 5); \
        else ENBL_ENT_TILE(T, 7);       \
    } while (0)
        if (itype == ENBL_U8) ENBL_ENT_N(uint8_t);
        else if (itype == ENBL_U16) ENBL_ENT_N(uint16_t);
        else ENBL_ENT_N(float);
#undef ENBL_ENT_N
#undef ENBL_ENT_TILE
        return 0;
    }
    if (n * n <= 49) {
        if (itype == ENBL_U8) hipLaunchKernelGGL((entropy_kernel<uint8_t, 49>), grid_for(w, h), dim3(kTile), 0, s, (const uint8_t*)img, ip, alpha, ap, w, h, n, tab, (uint8_t*)out, op);
        else if (itype == ENBL_U16) hipLaunchKernelGGL((entropy_kernel<uint16_t, 49>), grid_for(w, h), dim3(kTile), 0, s, (const uint16_t*)img, ip, alpha, ap, w, h, n, tab, (uint16_t*)out, op);
        else hipLaunchKernelGGL((entropy_kernel<float, 49>), grid_for(w, h), dim3(kTile), 0, s, (const float*)img, ip, alpha, ap, w, h, n, tab, (float*)out, op);
    } else {
        if (itype == ENBL_U8) hipLaunchKernelGGL((entropy_kernel<uint8_t, 121>), grid_for(w, h), dim3(kTile), 0, s, (const uint8_t*)img, ip, alpha, ap, w, h, n, tab, (uint8_t*)out, op);
        else if (itype == ENBL_U16) hipLaunchKernelGGL((entropy_kernel<uint16_t, 121>), grid_for(w, h), dim3(kTile), 0, s, (const uint16_t*)img, ip, alpha, ap, w, h, n, tab, (uint16_t*)out, op);
        else hipLaunchKernelGGL((entropy_kernel<float, 121>), grid_for(w, h), dim3(kTile), 0, s, (const float*)img, ip, alpha, ap, w, h, n, tab, (float*)out, op);
    }
    return 0;
}

int launch_entropy_weights(int itype, const void* img, long ip, const uint8_t* alpha, long ap, int w, int h, int n,
                           const double* tab, const WeightArgs& a, const float* lut, const void* grad, long gp,
                           float* mask, long mp, float* norm, long np_, uint8_t* ualpha, long up, hipStream_t s)
{
    if (n != 3 && n != 5 && n != 7) return ENBL_E_UNSUPPORTED;
    const dim3 grid((w + kEntTW - 1) / kEntTW, (h + kEntTH - 1) / kEntTH), block(kEntThreads);
#define ENBL_ENTW_TILE(T, N)                                                                                       \
    hipLaunchKernelGGL((entropy_tile_kernel<T, N, true>), grid, block, 0, s, (const T*)img, ip, alpha, ap, w, h, tab, \
                       (T*)nullptr, 0L,                                                                           \
                       FusedWeights<T>{a, lut, (const typename GradT<T>::type*)grad, gp, mask, mp, norm, np_, ualpha, up})
#define ENBL_ENTW_N(T)                  \
    do {                                \
        if (n == 3) ENBL_ENTW_TILE(T, 3); \
        else if (n == 5) ENBL_ENTW_TILE(T, 5); \
        else ENBL_ENTW_TILE(T, 7);       \
    } while (0)
    if (itype == ENBL_U8) ENBL_ENTW_N(uint8_t);
    else if (itype == ENBL_U16) ENBL_ENTW_N(uint16_t);
    else ENBL_ENTW_N(float);
#undef ENBL_ENTW_N
#undef ENBL_ENTW_TILE
    return 0;
}

int launch_weights(int itype, const void* img, long ip, const uint8_t* alpha, long ap, int w, int h,
                   const WeightArgs& a, const float* lut, const void* grad, long gp, const void* ent, long ep,
                   float* mask, long mp, float* norm, long np_, uint8_t* ualpha, long up, hipStream_t s)
{
    if (itype == ENBL_U8)
        hipLaunchKernelGGL((weights_kernel<uint8_t>), grid_for(w, h), dim3(kTile), 0, s, (const uint8_t*)img, ip, alpha, ap, w, a, lut,
                           (const int32_t*)grad, gp, (const uint8_t*)ent, ep, mask, mp, norm, np_, ualpha, up);
    else if (itype == ENBL_U16)
        hipLaunchKernelGGL((weights_kernel<uint16_t>), grid_for(w, h), dim3(kTile), 0, s, (const uint16_t*)img, ip, alpha, ap, w, a, lut,
                           (const int32_t*)grad, gp, (const uint16_t*)ent, ep, mask, mp, norm, np_, ualpha, up);
    else if (itype == ENBL_F32)
        hipLaunchKernelGGL((weights_kernel<float>), grid_for(w, h), dim3(kTile), 0, s, (const float*)img, ip, alpha, ap, w, a, lut,
                           (const float*)grad, gp, (const float*)ent, ep, mask, mp, norm, np_, ualpha, up);
    else return ENBL_E_UNSUPPORTED;
    return 0;
}

int launch_normalize(int mtype, const float* mask, long mp, const float* norm, long np_, int w, int h,
                     float fallback, double scale, int hard, void* out, long op, hipStream_t s)
{
    if (mtype == ENBL_I16) hipLaunchKernelGGL((normalize_kernel<int16_t>), grid_for(w, h), dim3(kTile), 0, s, mask, mp, norm, np_, w, fallback, scale, hard, (int16_t*)out, op);
    else if (mtype == ENBL_I32) hipLaunchKernelGGL((normalize_kernel<int32_t>), grid_for(w, h), dim3(kTile), 0, s, mask, mp, norm, np_, w, fallback, scale, hard, (int32_t*)out, op);
    else return ENBL_E_UNSUPPORTED;
    return 0;
}

int launch_hard_mask(float* masks, long mp, size_t plane, int w, int h, int K, hipStream_t s)
{
    hipLaunchKernelGGL(hard_mask_kernel, grid_for(w, h), dim3(kTile), 0, s, masks, mp, plane, w, K);
    return 0;
}

int launch_edge_rows(int itype, const EdgePlanes& e, const double* ka, int ra, const double* kb, int rb, hipStream_t s)
{
    const dim3 g((unsigned)((e.w + 255) / 256), (unsigned)e.h);
    if (itype == ENBL_F32) {
        if (kb) hipLaunchKernelGGL((edge_rows_kernel<float, float, true>), g, dim3(256), 0, s, e, ka, ra, kb, rb);
        else hipLaunchKernelGGL((edge_rows_kernel<float, float, false>), g, dim3(256), 0, s, e, ka, ra, kb, rb);
    } else {
        if (kb) hipLaunchKernelGGL((edge_rows_kernel<int32_t, double, true>), g, dim3(256), 0, s, e, ka, ra, kb, rb);
        else hipLaunchKernelGGL((edge_rows_kernel<int32_t, double, false>), g, dim3(256), 0, s, e, ka, ra, kb, rb);
    }
    return 0;
}

int launch_edge_sharpen(int itype, const EdgePlanes& e, const double* k, int r, double f, const void* gray, long gp,
                        void* out, long op, hipStream_t s)
{
    const dim3 g((unsigned)((e.w + 255) / 256), (unsigned)e.h);
    if (itype == ENBL_F32)
        hipLaunchKernelGGL((edge_sharpen_kernel<float, float>), g, dim3(256), 0, s, e, k, r, f, (const float*)gray, gp,
                           (float*)out, op);
    else
        hipLaunchKernelGGL((edge_sharpen_kernel<int32_t, double>), g, dim3(256), 0, s, e, k, r, f, (const int32_t*)gray,
                           gp, (int32_t*)out, op);
    return 0;
}

int launch_edge_log(int itype, const EdgePlanes& e, const double* ks, int rs, const double* kd, int rd,
                    const uint8_t* alpha, long ap, const void* lc, long lp, double mc, double smax, void* grad, long gp,
                    hipStream_t s)
{
    const dim3 g((unsigned)((e.w + 255) / 256), (unsigned)e.h);
    const bool fill = mc > 0.0;
    if (itype == ENBL_F32) {
        if (fill)
            hipLaunchKernelGGL((edge_log_kernel<float, float, true>), g, dim3(256), 0, s, e, ks, rs, kd, rd, alpha, ap,
                               (const float*)lc, lp, mc, smax, (float*)grad, gp);
        else
            hipLaunchKernelGGL((edge_log_kernel<float, float, false>), g, dim3(256), 0, s, e, ks, rs, kd, rd, alpha, ap,
                               (const float*)lc, lp, mc, smax, (float*)grad, gp);
    } else {
        if (fill)
            hipLaunchKernelGGL((edge_log_kernel<int32_t, double, true>), g, dim3(256), 0, s, e, ks, rs, kd, rd, alpha,
                               ap, (const int32_t*)lc, lp, mc, smax, (int32_t*)grad, gp);
        else
            hipLaunchKernelGGL((edge_log_kernel<int32_t, double, false>), g, dim3(256), 0, s, e, ks, rs, kd, rd, alpha,
                               ap, (const int32_t*)lc, lp, mc, smax, (int32_t*)grad, gp);
    }
    return 0;
}

int launch_entropy_clamp(int itype, const void* img, long ip, int w, int h, double lo, double up, void* out,
                         hipStream_t s)
{
    const dim3 g((unsigned)((w + kTile - 1) / kTile), (unsigned)h);
    if (itype == ENBL_U8)
        hipLaunchKernelGGL(entropy_clamp_kernel<uint8_t>, g, dim3(kTile), 0, s, (const uint8_t*)img, ip, w,
                           (uint8_t)lo, (uint8_t)up, (uint8_t)255, (uint8_t*)out);
    else if (itype == ENBL_U16)
        hipLaunchKernelGGL(entropy_clamp_kernel<uint16_t>, g, dim3(kTile), 0, s, (const uint16_t*)img, ip, w,
                           (uint16_t)lo, (uint16_t)up, (uint16_t)65535, (uint16_t*)out);
    else
        hipLaunchKernelGGL(entropy_clamp_kernel<float>, g, dim3(kTile), 0, s, (const float*)img, ip, w, (float)lo,
                           (float)up, FLT_MAX, (float*)out);
    return 0;
}

int launch_merge_many(int ptype, int mtype, const MergeMany& a, hipStream_t s)
{
    if (a.n < 1 || a.n > kMergeMax) return ENBL_E_ARG;
    const dim3 g((unsigned)((a.w + kTile - 1) / kTile), (unsigned)a.h);
    if (ptype == ENBL_I16 && mtype == ENBL_I16) { hipLaunchKernelGGL((merge_many_kernel<int16_t, int16_t>), g, dim3(kTile), 0, s, a); return 0; }
    if (ptype == ENBL_I32 && mtype == ENBL_I32) { hipLaunchKernelGGL((merge_many_kernel<int32_t, int32_t>), g, dim3(kTile), 0, s, a); return 0; }
    if (ptype == ENBL_F64 && mtype == ENBL_I32) { hipLaunchKernelGGL((merge_many_kernel<double, int32_t>), g, dim3(kTile), 0, s, a); return 0; }
    return ENBL_E_UNSUPPORTED;
}

}  // namespace enbl
