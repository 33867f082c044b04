// Blend colour spaces (--blend-colorspace=CIELAB|CIELUV; enfuse and enblend
// switch to CIELUV by themselves for integer images with an ICC profile,
// src/enfuse.cc:2380-2383), src/fixmath.h:429-1330.
//
// The reference converts every RGB pixel into the blend space on its way
// into the pyramid (ConvertVectorToLabPyramidFunctor,
// ConvertVectorToLuvPyramidFunctor) and back on its way out
// (ConvertLabPyramidToVectorFunctor, ConvertLuvPyramidToVectorFunctor):
//   in:  rgb = v / max;  XYZ (D50) = M * decode(rgb);  Lab (white D50) or
//        Luv (VIGRA's XYZ2Luv);  PyramidScale: L' = round(s L / 100),
//        a' = round(s (200 + a) / 400) with s = 2^(IntegerBits - 2); the
//        integer L', a', b' then enter the pyramid as an integer image would
//        (v << FractionBits);
//   out: the collapsed level through the integer image conversion (dither),
//        L = q 100 / s, a = q 400 / s - 200;  XYZ;  rgb = encode(M^-1 XYZ);
//        fromRealPromote(max * rgb), under the mask only.
// So the pyramid pipeline itself is unchanged: these two pointwise passes
// turn RGB images into "scaled blend-space" images of the same integer type
// and back.  The input profile is sRGB (the reference's assumption when an
// image carries no ICC profile): the IEC 61966-2-1 transfer curve as the
// parametric curve (a x + b)^2.4 / c x, primaries and D65 white adapted to
// the D50 connection space by the Bradford transform.  The arithmetic is
// double throughout.  Out-of-gamut results (a negative RGB component) are
// clamped by fromRealPromote; the reference polishes them with a numeric
// minimiser of the colour difference (OptimizableLuminanceSpace,
// src/fixmath.h:835-1045), which is not reproduced here.
#include <math.h>
#include <string.h>

#include "enbl_device.h"
#include "enbl_internal.h"

namespace enbl {
namespace {

constexpr int kCsTile = 256;

struct CsArgs {
    double m[9];     // linear RGB -> XYZ (D50)
    double mi[9];    // XYZ (D50) -> linear RGB
    double scale;    // 2^(IntegerBits - 2)
    double maxv;     // NumericTraits<T>::max()
    int space;       // ENBL_COLORSPACE_CIELAB / _CIELUV
};

// sRGB transfer curve as a parametric curve: (a x + b)^g for x >= d, else c x.
constexpr double kGamma = 2.4, kA = 1.0 / 1.055, kB = 0.055 / 1.055, kC = 1.0 / 12.92, kD = 0.04045;
// D50 white of the connection space and of the Lab profile.
constexpr double kXn = 0.9642, kYn = 1.0, kZn = 0.8249;
// VIGRA's XYZ2Luv / Luv2XYZ constants (white u', v' and the L* segments).
constexpr double kUn = 0.197839, kVn = 0.468342, kKappa = 24389.0 / 27.0, kEps = 216.0 / 24389.0;
constexpr double kColorLimit = 200.0;

__device__ __forceinline__ double decode(double x)
{
    return x >= kD ? pow(kA * x + kB, kGamma) : kC * x;
}

__device__ __forceinline__ double encode(double y)
{
    const double disc = pow(kA * kD + kB, kGamma);
    return y >= disc ? (pow(y, 1.0 / kGamma) - kB) / kA : y / kC;
}

__device__ __forceinline__ double lab_f(double t)
{
    const double limit = (24.0 / 116.0) * (24.0 / 116.0) * (24.0 / 116.0);
    return t <= limit ? (841.0 / 108.0) * t + (16.0 / 116.0) : cbrt(t);
}

__device__ __forceinline__ double lab_finv(double t)
{
    return t <= 24.0 / 116.0 ? (108.0 / 841.0) * (t - 16.0 / 116.0) : t * t * t;
}

template <class T> __device__ __forceinline__ T sat_scaled(double v, double maxv)
{
    // the rounded scaled value as the integer component the pyramid takes
    return (T)fmin(fmax(v, 0.0), maxv);
}

template <class T> __device__ __forceinline__ T from_real_comp(double v, double maxv)
{
    if (!(v >= 0.0)) return (T)0;
    return v > maxv ? (T)maxv : (T)(int32_t)(v + 0.5);
}

template <class T>
__global__ __launch_bounds__(kCsTile) void rgb_to_blend_kernel(CsArgs a, const T* __restrict__ img, long ip, int w,
                                                               T* __restrict__ out, long op)
{
    const int x = blockIdx.x * kCsTile + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const T* p = img + (long)y * ip + 3L * x;
    const double src = 1.0 / a.maxv;
    const double r = decode(src * (double)*gmem(p)), g = decode(src * (double)*gmem(p + 1)),
                 b = decode(src * (double)*gmem(p + 2));
    const double X = a.m[0] * r + a.m[1] * g + a.m[2] * b;
    const double Y = a.m[3] * r + a.m[4] * g + a.m[5] * b;
    const double Z = a.m[6] * r + a.m[7] * g + a.m[8] * b;
    double L, c1, c2;
    if (a.space == ENBL_COLORSPACE_CIELAB) {
        const double fx = lab_f(X / kXn), fy = lab_f(Y / kYn), fz = lab_f(Z / kZn);
        L = 116.0 * fy - 16.0;
        c1 = 500.0 * (fx - fy);
        c2 = 200.0 * (fy - fz);
    } else if (Y == 0.0) {
        L = c1 = c2 = 0.0;
    } else {
        L = Y < kEps ? kKappa * Y : 116.0 * pow(Y, 1.0 / 3.0) - 16.0;
        const double denom = X + 15.0 * Y + 3.0 * Z;
        c1 = 13.0 * L * (4.0 * X / denom - kUn);
        c2 = 13.0 * L * (9.0 * Y / denom - kVn);
    }
    T* q = out + (long)y * op + 3L * x;
    *gmem(q) = sat_scaled<T>(round(a.scale * L / 100.0), a.maxv);
    *gmem(q + 1) = sat_scaled<T>(round(a.scale * (kColorLimit + c1) / (2.0 * kColorLimit)), a.maxv);
    *gmem(q + 2) = sat_scaled<T>(round(a.scale * (kColorLimit + c2) / (2.0 * kColorLimit)), a.maxv);
}

template <class T>
__global__ __launch_bounds__(kCsTile) void blend_to_rgb_kernel(CsArgs a, const T* __restrict__ src, long sp,
                                                               const uint8_t* __restrict__ mask, long mp, int w,
                                                               T* __restrict__ img, long ip)
{
    const int x = blockIdx.x * kCsTile + threadIdx.x, y = blockIdx.y;
    if (x >= w || !*gmem(mask + (long)y * mp + x)) return;
    const T* s = src + (long)y * sp + 3L * x;
    const double L = (double)*gmem(s) * 100.0 / a.scale;
    const double c1 = (double)*gmem(s + 1) * (2.0 * kColorLimit) / a.scale - kColorLimit;
    const double c2 = (double)*gmem(s + 2) * (2.0 * kColorLimit) / a.scale - kColorLimit;
    double X, Y, Z;
    if (a.space == ENBL_COLORSPACE_CIELAB) {
        const double fy = (L + 16.0) / 116.0, fx = fy + c1 / 500.0, fz = fy - c2 / 200.0;
        X = kXn * lab_finv(fx);
        Y = kYn * lab_finv(fy);
        Z = kZn * lab_finv(fz);
    } else if (L == 0.0) {
        X = Y = Z = 0.0;
    } else {
        const double up = c1 / 13.0 / L + kUn, vp = c2 / 13.0 / L + kVn;
        Y = L < 8.0 ? L / kKappa : pow((L + 16.0) / 116.0, 3.0);
        X = 9.0 * up * Y / 4.0 / vp;
        Z = ((9.0 / vp - 15.0) * Y - X) / 3.0;
    }
    const double r = encode(a.mi[0] * X + a.mi[1] * Y + a.mi[2] * Z);
    const double g = encode(a.mi[3] * X + a.mi[4] * Y + a.mi[5] * Z);
    const double b = encode(a.mi[6] * X + a.mi[7] * Y + a.mi[8] * Z);
    T* q = img + (long)y * ip + 3L * x;
    *gmem(q) = from_real_comp<T>(a.maxv * r, a.maxv);
    *gmem(q + 1) = from_real_comp<T>(a.maxv * g, a.maxv);
    *gmem(q + 2) = from_real_comp<T>(a.maxv * b, a.maxv);
}

void mat_inverse(const double m[9], double r[9])
{
    const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                       m[2] * (m[3] * m[7] - m[4] * m[6]);
    r[0] = (m[4] * m[8] - m[5] * m[7]) / det;
    r[1] = (m[2] * m[7] - m[1] * m[8]) / det;
    r[2] = (m[1] * m[5] - m[2] * m[4]) / det;
    r[3] = (m[5] * m[6] - m[3] * m[8]) / det;
    r[4] = (m[0] * m[8] - m[2] * m[6]) / det;
    r[5] = (m[2] * m[3] - m[0] * m[5]) / det;
    r[6] = (m[3] * m[7] - m[4] * m[6]) / det;
    r[7] = (m[1] * m[6] - m[0] * m[7]) / det;
    r[8] = (m[0] * m[4] - m[1] * m[3]) / det;
}

void mat_mul(const double a[9], const double b[9], double r[9])
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}

// sRGB linear RGB -> XYZ relative to D50: the primaries' XYZ scaled so that
// RGB = 1 gives the D65 white, then the Bradford chromatic adaptation from
// D65 to D50 (the ICC connection space).
void srgb_to_xyz_d50(double m[9])
{
    const double xy[3][2] = {{0.64, 0.33}, {0.30, 0.60}, {0.15, 0.06}};
    const double wx = 0.3127, wy = 0.3290;
    double P[9];
    for (int k = 0; k < 3; ++k) {
        P[k] = xy[k][0] / xy[k][1];
        P[3 + k] = 1.0;
        P[6 + k] = (1.0 - xy[k][0] - xy[k][1]) / xy[k][1];
    }
    const double W[3] = {wx / wy, 1.0, (1.0 - wx - wy) / wy};
    double Pi[9];
    mat_inverse(P, Pi);
    double S[3];
    for (int i = 0; i < 3; ++i) S[i] = Pi[3 * i] * W[0] + Pi[3 * i + 1] * W[1] + Pi[3 * i + 2] * W[2];
    double M[9];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) M[3 * i + k] = P[3 * i + k] * S[k];
    const double B[9] = {0.8951, 0.2664, -0.1614, -0.7502, 1.7135, 0.0367, 0.0389, -0.0685, 1.0296};
    const double D50[3] = {kXn, kYn, kZn};
    double cs[3], cd[3];
    for (int i = 0; i < 3; ++i) {
        cs[i] = B[3 * i] * W[0] + B[3 * i + 1] * W[1] + B[3 * i + 2] * W[2];
        cd[i] = B[3 * i] * D50[0] + B[3 * i + 1] * D50[1] + B[3 * i + 2] * D50[2];
    }
    double Bi[9], DB[9], A[9];
    mat_inverse(B, Bi);
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) DB[3 * i + k] = cd[i] / cs[i] * B[3 * i + k];
    mat_mul(Bi, DB, A);
    mat_mul(A, M, m);
}

CsArgs cs_args(int space, int itype)
{
    CsArgs a;
    memset(&a, 0, sizeof a);
    srgb_to_xyz_d50(a.m);
    mat_inverse(a.m, a.mi);
    // PyramidScale: IntegerBits 9 for 8-bit, 17 for 16-bit image pyramids
    a.scale = itype == ENBL_U8 ? (double)(1 << 7) : (double)(1 << 15);
    a.maxv = itype == ENBL_U8 ? 255.0 : 65535.0;
    a.space = space;
    return a;
}

}  // namespace

namespace host {

int check_blend_space(int space, int itype, int channels)
{
    if (space == ENBL_COLORSPACE_IDENTITY || channels == 1) return ENBL_OK;
    if (space == ENBL_COLORSPACE_CIECAM)
        return fail(ENBL_E_UNSUPPORTED, "blend colour space CIECAM02 is not supported");
    if (space != ENBL_COLORSPACE_CIELAB && space != ENBL_COLORSPACE_CIELUV)
        return fail(ENBL_E_ARG, "unknown blend colour space %d", space);
    if (itype != ENBL_U8 && itype != ENBL_U16)
        return fail(ENBL_E_UNSUPPORTED, "blend colour spaces need 8- or 16-bit images");
    return ENBL_OK;
}

int rgb_to_blend_space(enbl_context* ctx, int space, int itype, int w, int h, const void* img, long ip, void* out,
                       long op)
{
    const CsArgs a = cs_args(space, itype);
    const dim3 grid((w + kCsTile - 1) / kCsTile, h);
    const double es = itype == ENBL_U8 ? 1.0 : 2.0;
    Timed t(ctx, "to_blend_space", (double)w * h * 6.0 * es);
    if (itype == ENBL_U8)
        hipLaunchKernelGGL(rgb_to_blend_kernel<uint8_t>, grid, dim3(kCsTile), 0, ctx->stream, a, (const uint8_t*)img,
                           ip, w, (uint8_t*)out, op);
    else
        hipLaunchKernelGGL(rgb_to_blend_kernel<uint16_t>, grid, dim3(kCsTile), 0, ctx->stream, a,
                           (const uint16_t*)img, ip, w, (uint16_t*)out, op);
    return launch_status("to_blend_space");
}

int blend_space_to_rgb_if(enbl_context* ctx, int space, int itype, int w, int h, const void* src, long sp,
                          const uint8_t* mask, long mp, void* img, long ip)
{
    const CsArgs a = cs_args(space, itype);
    const dim3 grid((w + kCsTile - 1) / kCsTile, h);
    const double es = itype == ENBL_U8 ? 1.0 : 2.0;
    Timed t(ctx, "from_blend_space", (double)w * h * (1.0 + 6.0 * es));
    if (itype == ENBL_U8)
        hipLaunchKernelGGL(blend_to_rgb_kernel<uint8_t>, grid, dim3(kCsTile), 0, ctx->stream, a, (const uint8_t*)src,
                           sp, mask, mp, w, (uint8_t*)img, ip);
    else
        hipLaunchKernelGGL(blend_to_rgb_kernel<uint16_t>, grid, dim3(kCsTile), 0, ctx->stream, a,
                           (const uint16_t*)src, sp, mask, mp, w, (uint16_t*)img, ip);
    return launch_status("from_blend_space");
}

}  // namespace host
}  // namespace enbl

using namespace enbl;
using namespace enbl::host;

extern "C" {

int enbl_set_blend_colorspace(enbl_context* ctx, int space)
{
    if (!ctx) return fail(ENBL_E_ARG, "enbl_set_blend_colorspace: null context");
    if (space < ENBL_COLORSPACE_IDENTITY || space > ENBL_COLORSPACE_CIECAM)
        return fail(ENBL_E_ARG, "unknown blend colour space %d", space);
    if (space == ENBL_COLORSPACE_CIECAM)
        return fail(ENBL_E_UNSUPPORTED, "blend colour space CIECAM02 is not supported");
    ctx->colorspace = space;
    return ENBL_OK;
}

int enbl_rgb_to_blend_space(enbl_context* ctx, int space, int itype, int w, int h, const void* img, long ipitch,
                            void* out, long opitch)
{
    if (!ctx || !img || !out) return fail(ENBL_E_ARG, "enbl_rgb_to_blend_space: null argument");
    if (w < 1 || h < 1 || ipitch < 3L * w || opitch < 3L * w) return fail(ENBL_E_SIZE, "enbl_rgb_to_blend_space: bad size or pitch");
    if (space == ENBL_COLORSPACE_IDENTITY) return fail(ENBL_E_ARG, "enbl_rgb_to_blend_space: identity space");
    int rc = check_blend_space(space, itype, 3);
    if (rc) return rc;
    return rgb_to_blend_space(ctx, space, itype, w, h, img, ipitch, out, opitch);
}

int enbl_blend_space_to_rgb_if(enbl_context* ctx, int space, int itype, int w, int h, const void* src, long spitch,
                               const uint8_t* mask, long mpitch, void* img, long ipitch)
{
    if (!ctx || !src || !mask || !img) return fail(ENBL_E_ARG, "enbl_blend_space_to_rgb_if: null argument");
    if (w < 1 || h < 1 || spitch < 3L * w || ipitch < 3L * w || mpitch < w)
        return fail(ENBL_E_SIZE, "enbl_blend_space_to_rgb_if: bad size or pitch");
    if (space == ENBL_COLORSPACE_IDENTITY) return fail(ENBL_E_ARG, "enbl_blend_space_to_rgb_if: identity space");
    int rc = check_blend_space(space, itype, 3);
    if (rc) return rc;
    return blend_space_to_rgb_if(ctx, space, itype, w, h, src, spitch, mask, mpitch, img, ipitch);
}

}  // extern "C"
