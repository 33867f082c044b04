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
template <class T, class R> struct Gray;
template <class R> struct Gray<uint8_t, R> {
    __device__ static R get(const uint8_t* p) { return (R)from_real_img<uint8_t>(((double)p[0] + (double)p[1] + (double)p[2]) / 3.0); }
};
template <class R> struct Gray<uint16_t, R> {
    __device__ static R get(const uint16_t* p) { return (R)from_real_img<uint16_t>(((double)p[0] + (double)p[1] + (double)p[2]) / 3.0); }
};
template <class R> struct Gray<float, R> {
    __device__ static R get(const float* p) { return (R)((double)(p[0] + p[1] + p[2]) / 3.0); }
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
__global__ __launch_bounds__(kTile) void gray_kernel(const T* __restrict__ img, long ip, int w, G* __restrict__ gray, long gp)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    gray[(long)y * gp + x] = Gray<T, G>::get(img + (long)y * ip + 3L * x);
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

template <class T> __device__ __forceinline__ float contrast_term(double wc, double g);
template <> __device__ __forceinline__ float contrast_term<uint8_t>(double wc, double g) { return (float)(wc * g / 255.0); }
template <> __device__ __forceinline__ float contrast_term<uint16_t>(double wc, double g) { return (float)(wc * g / 65535.0); }
template <> __device__ __forceinline__ float contrast_term<float>(double wc, double g) { return (float)(wc * g); }

template <class T> struct GradT { typedef int32_t type; };
template <> struct GradT<float> { typedef float type; };

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
        if (a.w_exposure > 0.0) {
            if (Img<T>::integral) {
                m = exp_lut[(int)Gray<T, T>::get(p)];
            } else {
                const float g = Gray<T, float>::get(p);
                const double yv = (double)(g / FLT_MAX);
                m = (float)(a.w_exposure * exposure_builtin(a.exposure_function, a.exposure_optimum, a.exposure_width, yv));
            }
        }
        if (a.w_contrast > 0.0) m = contrast_term<T>(a.w_contrast, (double)grad[(long)y * gp + x]) + m;
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
        if (a.w_entropy > 0.0) m = (float)(a.w_entropy * (double)(float)ent[(long)y * ep + x]) + m;
        ualpha[(long)y * up + x] = al;
    }
    mask[(long)y * mp + x] = m;
    float* nrm = norm + (long)y * np_ + x;
    *nrm = m + *nrm;
}

// Normalisation (src/enfuse.h:1529-1534) fused with the fixed-point
// conversion of mask level 0 (src/fixmath.h:188-194).
template <class M>
__global__ __launch_bounds__(kTile) void normalize_kernel(const float* __restrict__ mask, long mp,
                                                          const float* __restrict__ norm, long np_, int w,
                                                          float fallback, double scale, M* __restrict__ out, long op)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const float nv = norm[(long)y * np_ + x];
    const float v = nv > 0.0 ? (255.0f * mask[(long)y * mp + x]) / nv : fallback;
    out[(long)y * op + x] = from_real<M>((double)v * scale);
}

// ImageMaskMultiplyFunctor (src/enfuse.h:601-620) fused with the
// accumulation into the result pyramid (src/enfuse.h:1576-1588).
template <class P, class M, bool FIRST>
__global__ __launch_bounds__(kTile) void merge_kernel(const P* __restrict__ lp, long lpp, long lpl,
                                                      const M* __restrict__ g, long gpp, int w,
                                                      double maxpyr, P* __restrict__ res, long rp, long rl)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const double coeff = (double)g[(long)y * gpp + x] / maxpyr;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const P v = from_real<P>((double)lp[c * lpl + (long)y * lpp + x] * coeff);
        P* r = res + c * rl + (long)y * rp + x;
        if (FIRST) *r = v;
        else *r = Pix<P>::wrap(Pix<P>::A::add((typename Pix<P>::S)v, (typename Pix<P>::S)*r));
    }
}

}  // namespace

int launch_gray(int itype, const void* img, long ip, int w, int h, void* gray, long gp, hipStream_t s)
{
    if (itype == ENBL_U8) hipLaunchKernelGGL((gray_kernel<uint8_t, int32_t>), grid_for(w, h), dim3(kTile), 0, s, (const uint8_t*)img, ip, w, (int32_t*)gray, gp);
    else if (itype == ENBL_U16) hipLaunchKernelGGL((gray_kernel<uint16_t, int32_t>), grid_for(w, h), dim3(kTile), 0, s, (const uint16_t*)img, ip, w, (int32_t*)gray, gp);
    else if (itype == ENBL_F32) hipLaunchKernelGGL((gray_kernel<float, float>), grid_for(w, h), dim3(kTile), 0, s, (const float*)img, ip, w, (float*)gray, gp);
    else return ENBL_E_UNSUPPORTED;
    return 0;
}

int launch_stddev(int itype, const void* gray, long gp, const uint8_t* alpha, long ap, int w, int h, int n,
                  void* out, long op, hipStream_t s)
{
    if (itype == ENBL_F32) {
        if (n > 31) return ENBL_E_UNSUPPORTED;
        const int rows = h - 2 * (n / 2);
        if (rows <= 0) return 0;
        hipLaunchKernelGGL((stddev_float_rows_kernel<31>), dim3((rows + 63) / 64), dim3(64), 0, s,
                           (const float*)gray, gp, alpha, ap, w, h, n, (float*)out, op);
        return 0;
    }
    hipLaunchKernelGGL(stddev_int_kernel, grid_for(w, h), dim3(kTile), 0, s, (const int32_t*)gray, gp, alpha, ap, w, h, n,
                       (int32_t*)out, op);
    return 0;
}

int launch_entropy(int itype, const void* img, long ip, const uint8_t* alpha, long ap, int w, int h, int n,
                   const double* tab, void* out, long op, hipStream_t s)
{
    if (n * n > 121) return ENBL_E_UNSUPPORTED;
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
                     float fallback, double scale, void* out, long op, hipStream_t s)
{
    if (mtype == ENBL_I16) hipLaunchKernelGGL((normalize_kernel<int16_t>), grid_for(w, h), dim3(kTile), 0, s, mask, mp, norm, np_, w, fallback, scale, (int16_t*)out, op);
    else if (mtype == ENBL_I32) hipLaunchKernelGGL((normalize_kernel<int32_t>), grid_for(w, h), dim3(kTile), 0, s, mask, mp, norm, np_, w, fallback, scale, (int32_t*)out, op);
    else return ENBL_E_UNSUPPORTED;
    return 0;
}

int launch_merge(int ptype, int mtype, bool first, const void* lp, long lpp, long lpl, const void* g, long gpp,
                 int w, int h, double maxpyr, void* res, long rp, long rl, hipStream_t s)
{
#define ENBL_MERGE(P, M)                                                                                           \
    do {                                                                                                           \
        if (first) hipLaunchKernelGGL((merge_kernel<P, M, true>), grid_for(w, h), dim3(kTile), 0, s, (const P*)lp, lpp, lpl, \
                                      (const M*)g, gpp, w, maxpyr, (P*)res, rp, rl);                               \
        else hipLaunchKernelGGL((merge_kernel<P, M, false>), grid_for(w, h), dim3(kTile), 0, s, (const P*)lp, lpp, lpl,     \
                                (const M*)g, gpp, w, maxpyr, (P*)res, rp, rl);                                     \
        return 0;                                                                                                  \
    } while (0)
    if (ptype == ENBL_I16 && mtype == ENBL_I16) ENBL_MERGE(int16_t, int16_t);
    if (ptype == ENBL_I32 && mtype == ENBL_I32) ENBL_MERGE(int32_t, int32_t);
    if (ptype == ENBL_F64 && mtype == ENBL_I32) ENBL_MERGE(double, int32_t);
#undef ENBL_MERGE
    return ENBL_E_UNSUPPORTED;
}

}  // namespace enbl
