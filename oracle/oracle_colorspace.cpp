// CPU restatement of the blend colour spaces (TEST INFRASTRUCTURE ONLY:
// checker of the HIP colour-space path).
//
// src/fixmath.h:429-548 (PyramidScale), :622-681
// (ConvertVectorToLabPyramidFunctor), :1049-1130
// (ConvertLabPyramidToVectorFunctor), :1134-1205
// (ConvertVectorToLuvPyramidFunctor), :1214-1330
// (ConvertLuvPyramidToVectorFunctor), with the transforms enfuse and enblend
// build for images without an ICC profile (src/enfuse.cc:2385-2498): sRGB
// input profile, XYZ relative to D50, Lab with the profile's D50 media white,
// Luv by VIGRA's XYZ2Luv / Luv2XYZ functors.  Out-of-gamut polishing
// (OptimizableLuminanceSpace, src/fixmath.h:835-1045) is not restated: a
// negative RGB component is clamped by fromRealPromote.
//
// The colour mathematics here is written from the published definitions:
// IEC 61966-2-1 (the sRGB transfer curve and primaries, D65 white), the
// Bradford chromatic adaptation, CIE 1976 L*a*b* and L*u*v*, and the CIECAM02
// colour appearance model (CIE 159:2004) for the JCh space
// (ConvertVectorToJCHPyramidFunctor / ConvertJCHPyramidToVectorFunctor,
// src/fixmath.h:1336-1990) with the reference's viewing conditions
// (src/enblend.cc:2235-2244): white 100 x D50, Yb 20, La 31.83, average
// surround, D = 1.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>

#include "oracle.h"
#include "oracle_core.h"

namespace {

int g_space = ORC_COLORSPACE_IDENTITY;

const double kXn = 0.9642, kYn = 1.0, kZn = 0.8249;            // D50
const double kUn = 0.197839, kVn = 0.468342;                    // VIGRA's white u', v'
const double kKappa = 24389.0 / 27.0, kEps = 216.0 / 24389.0;  // CIE L* segments
const double kColorLimit = 200.0;

double srgb_decode(double x)
{
    // IEC 61966-2-1: ((x + 0.055) / 1.055)^2.4 above 0.04045, else x / 12.92,
    // written as the ICC parametric curve (a x + b)^g / c x
    const double a = 1.0 / 1.055, b = 0.055 / 1.055, c = 1.0 / 12.92;
    return x >= 0.04045 ? pow(a * x + b, 2.4) : c * x;
}

double srgb_encode(double y)
{
    const double a = 1.0 / 1.055, b = 0.055 / 1.055, c = 1.0 / 12.92;
    const double knee = pow(a * 0.04045 + b, 2.4);
    return y >= knee ? (pow(y, 1.0 / 2.4) - b) / a : y / c;
}

void inv3(const double m[9], double r[9])
{
    const double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
    const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    r[0] = c00 / det;
    r[1] = (m[2] * m[7] - m[1] * m[8]) / det;
    r[2] = (m[1] * m[5] - m[2] * m[4]) / det;
    r[3] = c01 / det;
    r[4] = (m[0] * m[8] - m[2] * m[6]) / det;
    r[5] = (m[2] * m[3] - m[0] * m[5]) / det;
    r[6] = c02 / det;
    r[7] = (m[1] * m[6] - m[0] * m[7]) / det;
    r[8] = (m[0] * m[4] - m[1] * m[3]) / det;
}

void mul3(const double a[9], const double b[9], double r[9])
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a[3 * i + k] * b[3 * k + j];
            r[3 * i + j] = s;
        }
}

// Linear sRGB -> XYZ relative to D50 (primaries scaled to the D65 white,
// then Bradford-adapted from D65 to D50).
void rgb_to_xyz_matrix(double out[9])
{
    const double px[3] = {0.64, 0.30, 0.15}, py[3] = {0.33, 0.60, 0.06};
    const double wx = 0.3127, wy = 0.3290;
    double P[9];
    for (int k = 0; k < 3; ++k) {
        P[0 + k] = px[k] / py[k];
        P[3 + k] = 1.0;
        P[6 + k] = (1.0 - px[k] - py[k]) / py[k];
    }
    const double W[3] = {wx / wy, 1.0, (1.0 - wx - wy) / wy};
    double Pi[9], S[3], M[9];
    inv3(P, Pi);
    for (int i = 0; i < 3; ++i) S[i] = Pi[3 * i] * W[0] + Pi[3 * i + 1] * W[1] + Pi[3 * i + 2] * W[2];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) M[3 * i + k] = P[3 * i + k] * S[k];
    const double Br[9] = {0.8951, 0.2664, -0.1614, -0.7502, 1.7135, 0.0367, 0.0389, -0.0685, 1.0296};
    const double D[3] = {kXn, kYn, kZn};
    double Bi[9], DB[9], A[9];
    inv3(Br, Bi);
    for (int i = 0; i < 3; ++i) {
        const double s = Br[3 * i] * W[0] + Br[3 * i + 1] * W[1] + Br[3 * i + 2] * W[2];
        const double d = Br[3 * i] * D[0] + Br[3 * i + 1] * D[1] + Br[3 * i + 2] * D[2];
        for (int k = 0; k < 3; ++k) DB[3 * i + k] = d / s * Br[3 * i + k];
    }
    mul3(Bi, DB, A);
    mul3(A, M, out);
}

double f_lab(double t)
{
    return t <= (24.0 / 116.0) * (24.0 / 116.0) * (24.0 / 116.0) ? (841.0 / 108.0) * t + 16.0 / 116.0 : cbrt(t);
}

double f_lab_inv(double t)
{
    return t <= 24.0 / 116.0 ? (108.0 / 841.0) * (t - 16.0 / 116.0) : t * t * t;
}

// RGB in [0, 1] -> (L, c1, c2) of the blend space.
void to_space(int space, const double M[9], const double rgb[3], double out[3])
{
    const double r = srgb_decode(rgb[0]), g = srgb_decode(rgb[1]), b = srgb_decode(rgb[2]);
    const double X = M[0] * r + M[1] * g + M[2] * b;
    const double Y = M[3] * r + M[4] * g + M[5] * b;
    const double Z = M[6] * r + M[7] * g + M[8] * b;
    if (space == ORC_COLORSPACE_CIELAB) {
        const double fx = f_lab(X / kXn), fy = f_lab(Y / kYn), fz = f_lab(Z / kZn);
        out[0] = 116.0 * fy - 16.0;
        out[1] = 500.0 * (fx - fy);
        out[2] = 200.0 * (fy - fz);
        return;
    }
    if (Y == 0.0) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    const double L = Y < kEps ? kKappa * Y : 116.0 * pow(Y, 1.0 / 3.0) - 16.0;
    const double den = X + 15.0 * Y + 3.0 * Z;
    out[0] = L;
    out[1] = 13.0 * L * (4.0 * X / den - kUn);
    out[2] = 13.0 * L * (9.0 * Y / den - kVn);
}

// (L, c1, c2) -> RGB (unclamped).
void from_space(int space, const double Mi[9], const double v[3], double rgb[3])
{
    double X, Y, Z;
    if (space == ORC_COLORSPACE_CIELAB) {
        const double fy = (v[0] + 16.0) / 116.0;
        X = kXn * f_lab_inv(fy + v[1] / 500.0);
        Y = kYn * f_lab_inv(fy);
        Z = kZn * f_lab_inv(fy - v[2] / 200.0);
    } else if (v[0] == 0.0) {
        X = Y = Z = 0.0;
    } else {
        const double up = v[1] / 13.0 / v[0] + kUn, vp = v[2] / 13.0 / v[0] + kVn;
        Y = v[0] < 8.0 ? v[0] / kKappa : pow((v[0] + 16.0) / 116.0, 3.0);
        X = 9.0 * up * Y / 4.0 / vp;
        Z = ((9.0 / vp - 15.0) * Y - X) / 3.0;
    }
    for (int i = 0; i < 3; ++i) rgb[i] = srgb_encode(Mi[3 * i] * X + Mi[3 * i + 1] * Y + Mi[3 * i + 2] * Z);
}

// ---- CIECAM02 (CIE 159:2004) ------------------------------------------------

// Published matrices: CAT02, its published inverse, Hunt-Pointer-Estevez and
// its published inverse -- as LittleCMS 2 (the library the reference calls,
// 2.12 here) evaluates them: its inverse transform takes the HPE inverse's
// entry (1, 2) as +0.000008 (published: -0.000008), and converts hue degrees
// to radians with pi = 3.141592654 in the inverse transform and in the
// eccentricity of the forward one.  The fixture written by LittleCMS itself
// (tests/golden/ciecam_lcms.npz) tells these apart: with them the forward
// model agrees to 1e-15 (J) and 1e-10 (C, h) relative, the inverse to 1e-14;
// with the published sign and the exact pi the inverse differs by 2e-5.
const double kCat02[9] = {0.7328, 0.4296, -0.1624, -0.7036, 1.6975, 0.0061, 0.0030, 0.0136, 0.9834};
const double kCat02Inv[9] = {1.096124, -0.278869, 0.182745, 0.454369, 0.473533, 0.072098,
                             -0.009628, -0.005698, 1.015326};
const double kHpe[9] = {0.38971, 0.68898, -0.07868, -0.22981, 1.18340, 0.04641, 0.0, 0.0, 1.0};
const double kHpeInv[9] = {1.910197, -1.112124, 0.201908, 0.370950, 0.629054, 0.000008, 0.0, 0.0, 1.0};
const double kCamD2R = 3.141592654 / 180.0;

struct Cam {
    double white[3];        // adopted white XYZ (Y = 100)
    double rgbw[3];         // its CAT02 responses
    double D, FL, n, z, Nbb, Ncb, c, Nc, Aw;
    double hpe_cat[9];      // M_HPE . M_CAT02^-1
    double cat_hpe[9];      // M_CAT02 . M_HPE^-1
};

void mat_vec(const double m[9], const double v[3], double r[3])
{
    for (int i = 0; i < 3; ++i) r[i] = m[3 * i] * v[0] + m[3 * i + 1] * v[1] + m[3 * i + 2] * v[2];
}

double compress(double FL, double x)
{
    // post-adaptation non-linear response compression
    const double t = pow(FL * fabs(x) / 100.0, 0.42);
    return (x < 0.0 ? -400.0 * t : 400.0 * t) / (t + 27.13) + 0.1;
}

double decompress(double FL, double y)
{
    const double d = y - 0.1, ad = fabs(d);
    const double v = 100.0 / FL * pow(27.13 * ad / (400.0 - ad), 1.0 / 0.42);
    return d < 0.0 ? -v : v;
}

Cam cam_setup()
{
    Cam m;
    m.white[0] = 100.0 * kXn;
    m.white[1] = 100.0 * kYn;
    m.white[2] = 100.0 * kZn;
    const double La = 31.83, Yb = 20.0;
    m.D = 1.0;
    m.c = 0.69;                      // average surround: F = 1.0, c = 0.69, Nc = 1.0
    m.Nc = 1.0;
    const double k = 1.0 / (5.0 * La + 1.0), k4 = k * k * k * k;
    m.FL = 0.2 * k4 * (5.0 * La) + 0.1 * (1.0 - k4) * (1.0 - k4) * pow(5.0 * La, 1.0 / 3.0);
    m.n = Yb / m.white[1];
    m.z = 1.48 + sqrt(m.n);
    m.Nbb = m.Ncb = 0.725 * pow(1.0 / m.n, 0.2);
    mul3(kHpe, kCat02Inv, m.hpe_cat);
    mul3(kCat02, kHpeInv, m.cat_hpe);
    mat_vec(kCat02, m.white, m.rgbw);
    double rc[3], rp[3], ra[3];
    for (int i = 0; i < 3; ++i) rc[i] = (m.white[1] * m.D / m.rgbw[i] + 1.0 - m.D) * m.rgbw[i];
    mat_vec(m.hpe_cat, rc, rp);
    for (int i = 0; i < 3; ++i) ra[i] = compress(m.FL, rp[i]);
    m.Aw = (2.0 * ra[0] + ra[1] + ra[2] / 20.0 - 0.305) * m.Nbb;
    return m;
}

// XYZ (Y of the white = 100) -> J, C, h (degrees in [0, 360)).
void cam_forward(const Cam& m, const double xyz[3], double jch[3])
{
    double rgb[3], rc[3], rp[3], ra[3];
    mat_vec(kCat02, xyz, rgb);
    for (int i = 0; i < 3; ++i) rc[i] = (m.white[1] * m.D / m.rgbw[i] + 1.0 - m.D) * rgb[i];
    mat_vec(m.hpe_cat, rc, rp);
    for (int i = 0; i < 3; ++i) ra[i] = compress(m.FL, rp[i]);
    const double a = ra[0] - 12.0 * ra[1] / 11.0 + ra[2] / 11.0;
    const double b = (ra[0] + ra[1] - 2.0 * ra[2]) / 9.0;
    double h = atan2(b, a) * (180.0 / M_PI);
    if (h < 0.0) h += 360.0;
    const double e = (12500.0 / 13.0) * m.Nc * m.Ncb * (cos(h * kCamD2R + 2.0) + 3.8);
    const double A = (2.0 * ra[0] + ra[1] + ra[2] / 20.0 - 0.305) * m.Nbb;
    const double J = 100.0 * pow(A / m.Aw, m.c * m.z);
    const double t = e * sqrt(a * a + b * b) / (ra[0] + ra[1] + (21.0 / 20.0) * ra[2]);
    const double C = pow(t, 0.9) * sqrt(J / 100.0) * pow(1.64 - pow(0.29, m.n), 0.73);
    jch[0] = J;
    jch[1] = C;
    jch[2] = h;
}

// J, C, h -> XYZ.
void cam_reverse(const Cam& m, const double jch[3], double xyz[3])
{
    const double J = jch[0], C = jch[1], h = jch[2];
    const double t = pow(C / (sqrt(J / 100.0) * pow(1.64 - pow(0.29, m.n), 0.73)), 1.0 / 0.9);
    const double hr = h * kCamD2R;
    const double e = (12500.0 / 13.0) * m.Nc * m.Ncb * (cos(hr + 2.0) + 3.8);
    const double A = m.Aw * pow(J / 100.0, 1.0 / (m.c * m.z));
    const double p1 = e / t, p2 = A / m.Nbb + 0.305, p3 = 21.0 / 20.0;
    const double sh = sin(hr), chh = cos(hr);
    double a, b;
    if (t == 0.0) {
        a = b = 0.0;
    } else if (fabs(sh) >= fabs(chh)) {
        const double p4 = p1 / sh;
        b = p2 * (2.0 + p3) * (460.0 / 1403.0) /
            (p4 + (2.0 + p3) * (220.0 / 1403.0) * (chh / sh) - 27.0 / 1403.0 + p3 * (6300.0 / 1403.0));
        a = b * (chh / sh);
    } else {
        const double p5 = p1 / chh;
        a = p2 * (2.0 + p3) * (460.0 / 1403.0) /
            (p5 + (2.0 + p3) * (220.0 / 1403.0) - (27.0 / 1403.0 - p3 * (6300.0 / 1403.0)) * (sh / chh));
        b = a * (sh / chh);
    }
    const double ra[3] = {(460.0 / 1403.0) * p2 + (451.0 / 1403.0) * a + (288.0 / 1403.0) * b,
                          (460.0 / 1403.0) * p2 - (891.0 / 1403.0) * a - (261.0 / 1403.0) * b,
                          (460.0 / 1403.0) * p2 - (220.0 / 1403.0) * a - (6300.0 / 1403.0) * b};
    double rp[3], rc[3], rgb[3];
    for (int i = 0; i < 3; ++i) rp[i] = decompress(m.FL, ra[i]);
    mat_vec(m.cat_hpe, rp, rc);
    for (int i = 0; i < 3; ++i) rgb[i] = rc[i] / (m.white[1] * m.D / m.rgbw[i] + 1.0 - m.D);
    mat_vec(kCat02Inv, rgb, xyz);
}

double wrap_cyclically(double x, double modulus)
{
    while (x < 0.0) x += modulus;
    return fmod(x, modulus);
}

// ConvertVectorToJCHPyramidFunctor (src/fixmath.h:1406-1464): RGB code ->
// (J s, C sin h s, C cos h s) as fixed point, s = 2^(IntegerBits - 8).
template <class T, class P>
void jch_to_level0(const T* img, long n, double maxv, double s, int frac, P* out, long plane)
{
    const Cam m = cam_setup();
    double M[9];
    rgb_to_xyz_matrix(M);
    const double unit = (double)(1 << frac);
#pragma omp parallel for schedule(guided)
    for (long i = 0; i < n; ++i) {
        double lin[3], xyz[3], jch[3];
        for (int c = 0; c < 3; ++c) lin[c] = srgb_decode(img[3 * i + c] / maxv);
        mat_vec(M, lin, xyz);
        for (int c = 0; c < 3; ++c) xyz[c] *= 100.0;
        cam_forward(m, xyz, jch);
        const double J = std::isnan(jch[0]) ? 0.0 : jch[0];
        const double C = std::isnan(jch[1]) ? 0.0 : jch[1];
        const double theta = wrap_cyclically(jch[2], 360.0) * (M_PI / 180.0);
        out[i] = orc::fromRealPromote<P>(J * s * unit);
        out[plane + i] = orc::fromRealPromote<P>(C * sin(theta) * s * unit);
        out[2 * plane + i] = orc::fromRealPromote<P>(C * cos(theta) * s * unit);
    }
}

// ConvertJCHPyramidToVectorFunctor (src/fixmath.h:1854-1977) without the
// out-of-gamut minimisers: J <= 0 gives black, other results are clamped.
template <class P, class T>
void level0_to_jch_rgb(const P* lv, long plane, long n, const uint8_t* mask, double maxv, double s, int frac, T* img)
{
    const Cam m = cam_setup();
    double M[9], Mi[9];
    rgb_to_xyz_matrix(M);
    inv3(M, Mi);
    const double unit = (double)(1 << frac);
#pragma omp parallel for schedule(guided)
    for (long i = 0; i < n; ++i) {
        if (!mask[i]) continue;
        const double j = (double)lv[i] / unit, cx = (double)lv[plane + i] / unit, cy = (double)lv[2 * plane + i] / unit;
        const double jch[3] = {j / s, hypot(cx / s, cy / s),
                               wrap_cyclically(atan2(cx / s, cy / s) * (180.0 / M_PI), 360.0)};
        double rgb[3] = {0.0, 0.0, 0.0};
        if (jch[0] > 0.0) {
            double xyz[3];
            cam_reverse(m, jch, xyz);
            for (int c = 0; c < 3; ++c) xyz[c] /= 100.0;
            for (int c = 0; c < 3; ++c) rgb[c] = srgb_encode(Mi[3 * c] * xyz[0] + Mi[3 * c + 1] * xyz[1] + Mi[3 * c + 2] * xyz[2]);
            for (int c = 0; c < 3; ++c) rgb[c] = std::min(std::max(rgb[c], 0.0), 1.0);
        }
        for (int c = 0; c < 3; ++c) {
            const double t = maxv * rgb[c];
            img[3 * i + c] = !(t >= 0.0) ? (T)0 : t > maxv ? (T)maxv : (T)(int32_t)(t + 0.5);
        }
    }
}

template <class T>
void to_scaled(int space, int w, int h, const T* img, T* out, double maxv, double s)
{
    double M[9];
    rgb_to_xyz_matrix(M);
    const double src = 1.0 / maxv;
#pragma omp parallel for schedule(guided)
    for (long i = 0; i < (long)w * h; ++i) {
        const double rgb[3] = {src * img[3 * i], src * img[3 * i + 1], src * img[3 * i + 2]};
        double v[3];
        to_space(space, M, rgb, v);
        // PyramidScale::scale_lightness_for_pyramid / _color_difference_
        const double q[3] = {round(s * v[0] / 100.0), round(s * (kColorLimit + v[1]) / (2.0 * kColorLimit)),
                             round(s * (kColorLimit + v[2]) / (2.0 * kColorLimit))};
        for (int c = 0; c < 3; ++c) out[3 * i + c] = (T)(q[c] < 0.0 ? 0.0 : q[c] > maxv ? maxv : q[c]);
    }
}

template <class T>
void from_scaled(int space, int w, int h, const T* src, const uint8_t* mask, T* img, double maxv, double s)
{
    double M[9], Mi[9];
    rgb_to_xyz_matrix(M);
    inv3(M, Mi);
#pragma omp parallel for schedule(guided)
    for (long i = 0; i < (long)w * h; ++i) {
        if (!mask[i]) continue;
        // PyramidScale::scale_lightness_of_pyramid / _color_difference_
        const double v[3] = {(double)src[3 * i] * 100.0 / s,
                             (double)src[3 * i + 1] * (2.0 * kColorLimit) / s - kColorLimit,
                             (double)src[3 * i + 2] * (2.0 * kColorLimit) / s - kColorLimit};
        double rgb[3];
        from_space(space, Mi, v, rgb);
        for (int c = 0; c < 3; ++c) {
            const double t = maxv * rgb[c];     // fromRealPromote
            img[3 * i + c] = !(t >= 0.0) ? (T)0 : t > maxv ? (T)maxv : (T)(int32_t)(t + 0.5);
        }
    }
}

}  // namespace

extern "C" {

int orc_set_blend_colorspace(int space)
{
    if (space != ORC_COLORSPACE_IDENTITY && space != ORC_COLORSPACE_CIELAB && space != ORC_COLORSPACE_CIELUV &&
        space != ORC_COLORSPACE_CIECAM)
        return -1;
    g_space = space;
    return 0;
}

int orc_blend_colorspace(void) { return g_space; }

int orc_rgb_to_blend_space(int space, int itype, int w, int h, const void* img, void* out)
{
    if (space != ORC_COLORSPACE_CIELAB && space != ORC_COLORSPACE_CIELUV) return -1;
    if (itype == ORC_U8) to_scaled<uint8_t>(space, w, h, (const uint8_t*)img, (uint8_t*)out, 255.0, 128.0);
    else if (itype == ORC_U16) to_scaled<uint16_t>(space, w, h, (const uint16_t*)img, (uint16_t*)out, 65535.0, 32768.0);
    else return -4;
    return 0;
}

int orc_blend_space_to_rgb_if(int space, int itype, int w, int h, const void* src, const uint8_t* mask, void* img)
{
    if (space != ORC_COLORSPACE_CIELAB && space != ORC_COLORSPACE_CIELUV) return -1;
    if (itype == ORC_U8) from_scaled<uint8_t>(space, w, h, (const uint8_t*)src, mask, (uint8_t*)img, 255.0, 128.0);
    else if (itype == ORC_U16)
        from_scaled<uint16_t>(space, w, h, (const uint16_t*)src, mask, (uint16_t*)img, 65535.0, 32768.0);
    else return -4;
    return 0;
}

// The blend space's (L, c1, c2) of RGB values in [0, 1] and back, unscaled
// (for pinning the colour mathematics itself).
int orc_colorspace_forward(int space, const double* rgb, long n, double* out)
{
    if (space != ORC_COLORSPACE_CIELAB && space != ORC_COLORSPACE_CIELUV) return -1;
    double M[9];
    rgb_to_xyz_matrix(M);
    for (long i = 0; i < n; ++i) to_space(space, M, rgb + 3 * i, out + 3 * i);
    return 0;
}

int orc_colorspace_inverse(int space, const double* v, long n, double* rgb)
{
    if (space != ORC_COLORSPACE_CIELAB && space != ORC_COLORSPACE_CIELUV) return -1;
    double M[9], Mi[9];
    rgb_to_xyz_matrix(M);
    inv3(M, Mi);
    for (long i = 0; i < n; ++i) from_space(space, Mi, v + 3 * i, rgb + 3 * i);
    return 0;
}

// The linear-RGB -> XYZ (D50) matrix, row major.
void orc_rgb_to_xyz_matrix(double* m) { rgb_to_xyz_matrix(m); }

// CIECAM02 under the reference's viewing conditions: XYZ (white Y = 100) to
// J, C, h and back.
int orc_ciecam_forward(const double* xyz, long n, double* jch)
{
    const Cam m = cam_setup();
    for (long i = 0; i < n; ++i) cam_forward(m, xyz + 3 * i, jch + 3 * i);
    return 0;
}

int orc_ciecam_reverse(const double* jch, long n, double* xyz)
{
    const Cam m = cam_setup();
    for (long i = 0; i < n; ++i) cam_reverse(m, jch + 3 * i, xyz + 3 * i);
    return 0;
}

// The JCh pyramid level 0 (planar, `plane` elements apart) of an 8- or
// 16-bit RGB image, and back under a mask.
int orc_rgb_to_jch_level0(int itype, int w, int h, const void* img, void* out, long plane)
{
    const long n = (long)w * h;
    if (itype == ORC_U8) jch_to_level0<uint8_t, int16_t>((const uint8_t*)img, n, 255.0, 2.0, 7, (int16_t*)out, plane);
    else if (itype == ORC_U16) jch_to_level0<uint16_t, int32_t>((const uint16_t*)img, n, 65535.0, 512.0, 7, (int32_t*)out, plane);
    else return -4;
    return 0;
}

int orc_jch_level0_to_rgb_if(int itype, int w, int h, const void* lv, long plane, const uint8_t* mask, void* img)
{
    const long n = (long)w * h;
    if (itype == ORC_U8) level0_to_jch_rgb<int16_t, uint8_t>((const int16_t*)lv, plane, n, mask, 255.0, 2.0, 7, (uint8_t*)img);
    else if (itype == ORC_U16) level0_to_jch_rgb<int32_t, uint16_t>((const int32_t*)lv, plane, n, mask, 65535.0, 512.0, 7, (uint16_t*)img);
    else return -4;
    return 0;
}

}  // extern "C"
