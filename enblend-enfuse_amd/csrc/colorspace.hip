// Blend colour spaces (--blend-colorspace=CIELAB|CIELUV|CIECAM; enfuse and enblend
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

// ---- CIECAM02 JCh (--blend-colorspace=CIECAM, src/fixmath.h:1336-1990) ----
//
// in:  rgb = v / max;  XYZ (D50) as above, times 100;  CIECAM02 forward
//      (CIE 159:2004) under the reference's viewing conditions
//      (src/enblend.cc:2235-2244: white 100 x D50, Yb 20, La 31.83, average
//      surround, D = 1) -> J, C, h;  NaN J or C -> 0;  the pyramid's level 0
//      is fromRealPromote(2^frac x (J s, C sin h s, C cos h s)) with
//      s = 2^(IntegerBits - 8) (PyramidScaleJCh): fixed point with a
//      fraction, written straight into the pyramid;
// out: level 0 / 2^frac (no dither), J = j / s, C = hypot, h = atan2 in
//      degrees wrapped to [0, 360);  J <= 0 gives black;  CIECAM02 inverse,
//      XYZ / 100, linear RGB, the sRGB curve, clamped to [0, 1] (the
//      reference polishes out-of-gamut colours with GSL minimisers, not
//      reproduced), fromRealPromote(max rgb) under the mask.
// The model's matrices and the degree conversions are those LittleCMS 2
// (which the reference calls) evaluates: the HPE inverse's entry (1, 2) as
// +0.000008 and pi = 3.141592654 in its inverse transform and in the forward
// eccentricity.
struct CamArgs {
    double m[9], mi[9];        // linear RGB <-> XYZ (D50)
    double cat[9], cat_inv[9]; // CAT02 and its published inverse
    double hpe_cat[9];         // M_HPE . M_CAT02^-1
    double cat_hpe[9];         // M_CAT02 . M_HPE^-1
    double white_y, rgbw[3], D, FL, n, z, Nbb, Ncb, c, Nc, Aw, cnorm;
    double s, unit, maxv;      // PyramidScaleJCh scale, 2^frac, image maximum
};

constexpr double kCamD2R = 3.141592654 / 180.0;

__host__ __device__ inline void mvec(const double m[9], double x, double y, double z, double r[3])
{
    r[0] = m[0] * x + m[1] * y + m[2] * z;
    r[1] = m[3] * x + m[4] * y + m[5] * z;
    r[2] = m[6] * x + m[7] * y + m[8] * z;
}

__host__ __device__ inline double cam_compress(double FL, double x)
{
    const double t = pow(FL * fabs(x) / 100.0, 0.42);
    return (x < 0.0 ? -400.0 * t : 400.0 * t) / (t + 27.13) + 0.1;
}

__device__ inline double cam_decompress(double FL, double y)
{
    const double d = y - 0.1, ad = fabs(d);
    const double v = 100.0 / FL * pow(27.13 * ad / (400.0 - ad), 1.0 / 0.42);
    return d < 0.0 ? -v : v;
}

__device__ inline double wrap360(double x)
{
    while (x < 0.0) x += 360.0;
    return fmod(x, 360.0);
}

template <class T, class P>
__global__ __launch_bounds__(kCsTile) void rgb_to_jch_kernel(CamArgs a, const T* __restrict__ img, long ip, int w,
                                                             P* __restrict__ out, long op, long opl)
{
    const int x = blockIdx.x * kCsTile + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const T* p = img + (long)y * ip + 3L * x;
    const double src = 1.0 / a.maxv;
    const double r = decode(src * (double)*gmem(p)), g = decode(src * (double)*gmem(p + 1)),
                 b = decode(src * (double)*gmem(p + 2));
    double xyz[3], rgb[3], rp[3];
    mvec(a.m, r, g, b, xyz);
    mvec(a.cat, 100.0 * xyz[0], 100.0 * xyz[1], 100.0 * xyz[2], rgb);
    double rc[3];
    for (int i = 0; i < 3; ++i) rc[i] = (a.white_y * a.D / a.rgbw[i] + 1.0 - a.D) * rgb[i];
    mvec(a.hpe_cat, rc[0], rc[1], rc[2], rp);
    const double r0 = cam_compress(a.FL, rp[0]), r1 = cam_compress(a.FL, rp[1]), r2 = cam_compress(a.FL, rp[2]);
    const double ca = r0 - 12.0 * r1 / 11.0 + r2 / 11.0;
    const double cb = (r0 + r1 - 2.0 * r2) / 9.0;
    double h = atan2(cb, ca) * (180.0 / M_PI);
    if (h < 0.0) h += 360.0;
    const double e = (12500.0 / 13.0) * a.Nc * a.Ncb * (cos(h * kCamD2R + 2.0) + 3.8);
    const double A = (2.0 * r0 + r1 + r2 / 20.0 - 0.305) * a.Nbb;
    double J = 100.0 * pow(A / a.Aw, a.c * a.z);
    const double t = e * sqrt(ca * ca + cb * cb) / (r0 + r1 + (21.0 / 20.0) * r2);
    double C = pow(t, 0.9) * sqrt(J / 100.0) * a.cnorm;
    if (isnan(J)) J = 0.0;
    if (isnan(C)) C = 0.0;
    const double theta = wrap360(h) * (M_PI / 180.0);
    P* q = out + (long)y * op + x;
    *gmem(q) = from_real<P>(J * a.s * a.unit);
    *gmem(q + opl) = from_real<P>(C * sin(theta) * a.s * a.unit);
    *gmem(q + 2 * opl) = from_real<P>(C * cos(theta) * a.s * a.unit);
}

template <class P, class T>
__global__ __launch_bounds__(kCsTile) void jch_to_rgb_kernel(CamArgs a, const P* __restrict__ lv, long lp, long lpl,
                                                             const uint8_t* __restrict__ mask, long mp, int w,
                                                             T* __restrict__ img, long ip)
{
    const int x = blockIdx.x * kCsTile + threadIdx.x, y = blockIdx.y;
    if (x >= w || !*gmem(mask + (long)y * mp + x)) return;
    const P* q = lv + (long)y * lp + x;
    const double j = (double)*gmem(q) / a.unit, cx = (double)*gmem(q + lpl) / a.unit,
                 cy = (double)*gmem(q + 2 * lpl) / a.unit;
    const double J = j / a.s, C = hypot(cx / a.s, cy / a.s);
    const double h = wrap360(atan2(cx / a.s, cy / a.s) * (180.0 / M_PI));
    double out[3] = {0.0, 0.0, 0.0};
    if (J > 0.0) {
        const double t = pow(C / (sqrt(J / 100.0) * a.cnorm), 1.0 / 0.9);
        const double hr = h * kCamD2R;
        const double e = (12500.0 / 13.0) * a.Nc * a.Ncb * (cos(hr + 2.0) + 3.8);
        const double A = a.Aw * pow(J / 100.0, 1.0 / (a.c * a.z));
        const double p1 = e / t, p2 = A / a.Nbb + 0.305, p3 = 21.0 / 20.0;
        const double sh = sin(hr), chh = cos(hr);
        double ca, cb;
        if (t == 0.0) {
            ca = cb = 0.0;
        } else if (fabs(sh) >= fabs(chh)) {
            const double p4 = p1 / sh;
            cb = p2 * (2.0 + p3) * (460.0 / 1403.0) /
                 (p4 + (2.0 + p3) * (220.0 / 1403.0) * (chh / sh) - 27.0 / 1403.0 + p3 * (6300.0 / 1403.0));
            ca = cb * (chh / sh);
        } else {
            const double p5 = p1 / chh;
            ca = p2 * (2.0 + p3) * (460.0 / 1403.0) /
                 (p5 + (2.0 + p3) * (220.0 / 1403.0) - (27.0 / 1403.0 - p3 * (6300.0 / 1403.0)) * (sh / chh));
            cb = ca * (sh / chh);
        }
        const double r0 = (460.0 / 1403.0) * p2 + (451.0 / 1403.0) * ca + (288.0 / 1403.0) * cb;
        const double r1 = (460.0 / 1403.0) * p2 - (891.0 / 1403.0) * ca - (261.0 / 1403.0) * cb;
        const double r2 = (460.0 / 1403.0) * p2 - (220.0 / 1403.0) * ca - (6300.0 / 1403.0) * cb;
        double rc[3], rgb[3], xyz[3];
        mvec(a.cat_hpe, cam_decompress(a.FL, r0), cam_decompress(a.FL, r1), cam_decompress(a.FL, r2), rc);
        for (int i = 0; i < 3; ++i) rgb[i] = rc[i] / (a.white_y * a.D / a.rgbw[i] + 1.0 - a.D);
        mvec(a.cat_inv, rgb[0], rgb[1], rgb[2], xyz);
        for (int i = 0; i < 3; ++i) xyz[i] /= 100.0;
        double lin[3];
        mvec(a.mi, xyz[0], xyz[1], xyz[2], lin);
        for (int i = 0; i < 3; ++i) out[i] = fmin(fmax(encode(lin[i]), 0.0), 1.0);
    }
    T* o = img + (long)y * ip + 3L * x;
    *gmem(o) = from_real_comp<T>(a.maxv * out[0], a.maxv);
    *gmem(o + 1) = from_real_comp<T>(a.maxv * out[1], a.maxv);
    *gmem(o + 2) = from_real_comp<T>(a.maxv * out[2], a.maxv);
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

CamArgs cam_args(int itype)
{
    CamArgs a;
    memset(&a, 0, sizeof a);
    srgb_to_xyz_d50(a.m);
    mat_inverse(a.m, a.mi);
    const double cat[9] = {0.7328, 0.4296, -0.1624, -0.7036, 1.6975, 0.0061, 0.0030, 0.0136, 0.9834};
    const double cat_inv[9] = {1.096124, -0.278869, 0.182745, 0.454369, 0.473533, 0.072098,
                               -0.009628, -0.005698, 1.015326};
    const double hpe[9] = {0.38971, 0.68898, -0.07868, -0.22981, 1.18340, 0.04641, 0.0, 0.0, 1.0};
    const double hpe_inv[9] = {1.910197, -1.112124, 0.201908, 0.370950, 0.629054, 0.000008, 0.0, 0.0, 1.0};
    memcpy(a.cat, cat, sizeof cat);
    memcpy(a.cat_inv, cat_inv, sizeof cat_inv);
    mat_mul(hpe, cat_inv, a.hpe_cat);
    mat_mul(cat, hpe_inv, a.cat_hpe);
    const double white[3] = {100.0 * kXn, 100.0 * kYn, 100.0 * kZn};
    const double La = 31.83, Yb = 20.0;
    a.white_y = white[1];
    a.D = 1.0;
    a.c = 0.69;   // average surround: F = 1.0, c = 0.69, Nc = 1.0
    a.Nc = 1.0;
    const double k = 1.0 / (5.0 * La + 1.0), k4 = k * k * k * k;
    a.FL = 0.2 * k4 * (5.0 * La) + 0.1 * (1.0 - k4) * (1.0 - k4) * pow(5.0 * La, 1.0 / 3.0);
    a.n = Yb / white[1];
    a.z = 1.48 + sqrt(a.n);
    a.Nbb = a.Ncb = 0.725 * pow(1.0 / a.n, 0.2);
    a.cnorm = pow(1.64 - pow(0.29, a.n), 0.73);
    mvec(a.cat, white[0], white[1], white[2], a.rgbw);
    double rc[3], rp[3];
    for (int i = 0; i < 3; ++i) rc[i] = (a.white_y * a.D / a.rgbw[i] + 1.0 - a.D) * a.rgbw[i];
    mvec(a.hpe_cat, rc[0], rc[1], rc[2], rp);
    const double r0 = cam_compress(a.FL, rp[0]), r1 = cam_compress(a.FL, rp[1]), r2 = cam_compress(a.FL, rp[2]);
    a.Aw = (2.0 * r0 + r1 + r2 / 20.0 - 0.305) * a.Nbb;
    // PyramidScaleJCh: 2^(IntegerBits - 1 - 7), IntegerBits 9 (8-bit) / 17 (16-bit)
    a.s = itype == ENBL_U8 ? 2.0 : 512.0;
    a.unit = 128.0;
    a.maxv = itype == ENBL_U8 ? 255.0 : 65535.0;
    return a;
}

}  // namespace

namespace host {

int check_blend_space(int space, int itype, int channels)
{
    if (space == ENBL_COLORSPACE_IDENTITY || channels == 1) return ENBL_OK;
    if (space != ENBL_COLORSPACE_CIELAB && space != ENBL_COLORSPACE_CIELUV && space != ENBL_COLORSPACE_CIECAM)
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

int rgb_to_jch(enbl_context* ctx, int itype, int w, int h, const void* img, long ip, void* out, long op, long opl)
{
    const CamArgs a = cam_args(itype);
    const dim3 grid((w + kCsTile - 1) / kCsTile, h);
    const double es = itype == ENBL_U8 ? 1.0 : 2.0;
    Timed t(ctx, "to_jch", (double)w * h * 3.0 * (es + 2.0 * es));
    if (itype == ENBL_U8)
        hipLaunchKernelGGL((rgb_to_jch_kernel<uint8_t, int16_t>), grid, dim3(kCsTile), 0, ctx->stream, a,
                           (const uint8_t*)img, ip, w, (int16_t*)out, op, opl);
    else
        hipLaunchKernelGGL((rgb_to_jch_kernel<uint16_t, int32_t>), grid, dim3(kCsTile), 0, ctx->stream, a,
                           (const uint16_t*)img, ip, w, (int32_t*)out, op, opl);
    return launch_status("to_jch");
}

int jch_to_rgb_if(enbl_context* ctx, int itype, int w, int h, const void* lv, long lp, long lpl, const uint8_t* mask,
                  long mp, void* img, long ip)
{
    const CamArgs a = cam_args(itype);
    const dim3 grid((w + kCsTile - 1) / kCsTile, h);
    const double es = itype == ENBL_U8 ? 1.0 : 2.0;
    Timed t(ctx, "from_jch", (double)w * h * (1.0 + 3.0 * (2.0 * es + es)));
    if (itype == ENBL_U8)
        hipLaunchKernelGGL((jch_to_rgb_kernel<int16_t, uint8_t>), grid, dim3(kCsTile), 0, ctx->stream, a,
                           (const int16_t*)lv, lp, lpl, mask, mp, w, (uint8_t*)img, ip);
    else
        hipLaunchKernelGGL((jch_to_rgb_kernel<int32_t, uint16_t>), grid, dim3(kCsTile), 0, ctx->stream, a,
                           (const int32_t*)lv, lp, lpl, mask, mp, w, (uint16_t*)img, ip);
    return launch_status("from_jch");
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
    ctx->colorspace = space;
    return ENBL_OK;
}

int enbl_rgb_to_blend_space(enbl_context* ctx, int space, int itype, int w, int h, const void* img, long ipitch,
                            void* out, long opitch)
{
    if (!ctx || !img || !out) return fail(ENBL_E_ARG, "enbl_rgb_to_blend_space: null argument");
    if (w < 1 || h < 1 || ipitch < 3L * w || opitch < 3L * w) return fail(ENBL_E_SIZE, "enbl_rgb_to_blend_space: bad size or pitch");
    if (space == ENBL_COLORSPACE_IDENTITY) return fail(ENBL_E_ARG, "enbl_rgb_to_blend_space: identity space");
    if (space == ENBL_COLORSPACE_CIECAM)
        return fail(ENBL_E_ARG, "enbl_rgb_to_blend_space: CIECAM02 converts into a pyramid level (enbl_rgb_to_jch)");
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
    if (space == ENBL_COLORSPACE_CIECAM)
        return fail(ENBL_E_ARG, "enbl_blend_space_to_rgb_if: CIECAM02 converts from a pyramid level (enbl_jch_to_rgb_if)");
    int rc = check_blend_space(space, itype, 3);
    if (rc) return rc;
    return blend_space_to_rgb_if(ctx, space, itype, w, h, src, spitch, mask, mpitch, img, ipitch);
}

int enbl_rgb_to_jch(enbl_context* ctx, int itype, int w, int h, const void* img, long ipitch, void* out, long opitch,
                    long oplane)
{
    if (!ctx || !img || !out) return fail(ENBL_E_ARG, "enbl_rgb_to_jch: null argument");
    if (w < 1 || h < 1 || ipitch < 3L * w || opitch < w || oplane < opitch * h)
        return fail(ENBL_E_SIZE, "enbl_rgb_to_jch: bad size, pitch or plane");
    if (itype != ENBL_U8 && itype != ENBL_U16) return fail(ENBL_E_UNSUPPORTED, "CIECAM02 needs 8- or 16-bit images");
    return rgb_to_jch(ctx, itype, w, h, img, ipitch, out, opitch, oplane);
}

int enbl_jch_to_rgb_if(enbl_context* ctx, int itype, int w, int h, const void* lv, long lpitch, long lplane,
                       const uint8_t* mask, long mpitch, void* img, long ipitch)
{
    if (!ctx || !lv || !mask || !img) return fail(ENBL_E_ARG, "enbl_jch_to_rgb_if: null argument");
    if (w < 1 || h < 1 || lpitch < w || lplane < lpitch * h || mpitch < w || ipitch < 3L * w)
        return fail(ENBL_E_SIZE, "enbl_jch_to_rgb_if: bad size, pitch or plane");
    if (itype != ENBL_U8 && itype != ENBL_U16) return fail(ENBL_E_UNSUPPORTED, "CIECAM02 needs 8- or 16-bit images");
    return jch_to_rgb_if(ctx, itype, w, h, lv, lpitch, lplane, mask, mpitch, img, ipitch);
}

}  // extern "C"
