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
// Bradford chromatic adaptation, CIE 1976 L*a*b* and L*u*v*.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

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
    if (space != ORC_COLORSPACE_IDENTITY && space != ORC_COLORSPACE_CIELAB && space != ORC_COLORSPACE_CIELUV)
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

}  // extern "C"
