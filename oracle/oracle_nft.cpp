// FP32 restatement of enblend's nearest-feature-transform seam mask
// (TEST INFRASTRUCTURE ONLY: checker of the HIP NFT path).
//
//   nearestFeatureTransform, open boundaries (src/nearest.h:424-552):
//     diff12 = saturating white - black alpha, diff21 = black - white
//     (src/nearest.h:359-372, :464-468, :494-498); overlap tally against
//     2 * 2 * (w + h) (src/nearest.h:449-457, :523-538);
//     dist = distance transform with background 0; mask = dist12 < dist21
//     ? 255 : 0 (src/nearest.h:540-545).
//   Distance transform of the OpenMP build (src/openmp_vigra.h:244-470):
//     Felzenszwalb-Huttenlocher 1-D lower-envelope transform in float over
//     every column (f = FLT_MAX on background, 0 on features), then over
//     every row of the intermediate image; sqrt of the result for the
//     Euclidean norm.  Every float operation is kept in the reference's
//     order, so the float roundings (q^2 beyond 2^24, the envelope
//     intersections, the final sum and sqrt) are the reference's.
#include <float.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "oracle.h"

namespace {

// ManhattanTransform1D (src/openmp_vigra.h:283-301).
void manhattan_1d(float* d, const float* f, int n)
{
    d[0] = f[0];
    for (int q = 1; q < n; ++q) d[q] = std::min<float>(f[q], d[q - 1] + 1.0f);
    for (int q = n - 2; q >= 0; --q) d[q] = std::min<float>(d[q], d[q + 1] + 1.0f);
}

// Intersection of the parabolas of q and v (src/openmp_vigra.h:322, :328):
// (sum_q - (f[v] + v^2)) / (2 (q - v)).  PLANTED = 1 is a deliberately wrong
// operation order, ((sum_q - f[v]) - v^2) / (2 (q - v)), reachable only
// through orc_euclid_1d so that tests can show they detect it.
template <int PLANTED>
inline float intersection(float sum_q, float fv, int v, int q)
{
    if (PLANTED == 1) return ((sum_q - fv) - (float)(v * v)) / (float)(2 * (q - v));
    return (sum_q - (fv + (float)(v * v))) / (float)(2 * (q - v));
}

// EuclideanTransform1D (src/openmp_vigra.h:304-357): v, z scratch of n and
// n + 1 entries.
template <int PLANTED = 0>
void euclid_1d(float* d, const float* f, int n, int* v, float* z)
{
    const float inf = INFINITY;
    int k = 0;
    v[0] = 0;
    z[0] = -inf;
    z[1] = inf;
    for (int q = 1; q < n; ++q) {
        const float sum_q = f[q] + (float)q * (float)q;
        float s = intersection<PLANTED>(sum_q, f[v[k]], v[k], q);
        while (s <= z[k]) {
            --k;
            s = intersection<PLANTED>(sum_q, f[v[k]], v[k], q);
        }
        ++k;
        v[k] = q;
        z[k] = s;
        z[k + 1] = inf;
    }
    k = 0;
    for (int q = 0; q < n; ++q) {
        while (z[k + 1] < (float)q) ++k;
        d[q] = (float)((q - v[k]) * (q - v[k])) + f[v[k]];
    }
}

// fhDistanceTransform (src/openmp_vigra.h:360-438): feature = src != 0.
void fh_transform(const uint8_t* src, long spitch, int w, int h, int norm, float* dst, long dpitch)
{
    std::vector<float> inter((size_t)w * h);
    const int L = std::max(w, h);
#pragma omp parallel
    {
        std::vector<float> f(L), d(L), z(L + 1);
        std::vector<int> v(L);
#pragma omp for schedule(guided)
        for (int x = 0; x < w; ++x) {
            for (int y = 0; y < h; ++y) f[y] = src[(long)y * spitch + x] == 0 ? FLT_MAX : 0.0f;
            if (norm == 1) manhattan_1d(d.data(), f.data(), h);
            else euclid_1d(d.data(), f.data(), h, v.data(), z.data());
            for (int y = 0; y < h; ++y) inter[(size_t)y * w + x] = d[y];
        }
#pragma omp for schedule(guided)
        for (int y = 0; y < h; ++y) {
            const float* row = &inter[(size_t)y * w];
            if (norm == 1) manhattan_1d(d.data(), row, w);
            else euclid_1d(d.data(), row, w, v.data(), z.data());
            for (int x = 0; x < w; ++x) dst[(long)y * dpitch + x] = norm == 1 ? d[x] : sqrtf(d[x]);
        }
    }
}

}  // namespace

extern "C" {

int orc_euclid_1d(const float* f, int n, int planted, float* d)
{
    if (n < 1 || planted < 0 || planted > 1) return -1;
    std::vector<int> v((size_t)n);
    std::vector<float> z((size_t)n + 1);
    if (planted) euclid_1d<1>(d, f, n, v.data(), z.data());
    else euclid_1d<0>(d, f, n, v.data(), z.data());
    return 0;
}

int orc_distance_transform(const uint8_t* src, long spitch, int w, int h, int norm, float* dst, long dpitch)
{
    if (w < 1 || h < 1 || (norm != 1 && norm != 2)) return -1;
    fh_transform(src, spitch, w, h, norm, dst, dpitch);
    return 0;
}

int orc_nft_mask(const uint8_t* white, long wpitch, const uint8_t* black, long bpitch, int w, int h, int norm,
                 uint8_t* mask, long mpitch)
{
    if (w < 1 || h < 1 || (norm != 1 && norm != 2)) return -1;
    std::vector<uint8_t> d12((size_t)w * h), d21((size_t)w * h);
    long n12 = 0, n21 = 0;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const uint8_t a = white[(long)y * wpitch + x], b = black[(long)y * bpitch + x];
            d12[(size_t)y * w + x] = b < a ? (uint8_t)(a - b) : 0;
            d21[(size_t)y * w + x] = a < b ? (uint8_t)(b - a) : 0;
            n12 += d12[(size_t)y * w + x] != 0;
            n21 += d21[(size_t)y * w + x] != 0;
        }
    std::vector<float> t12((size_t)w * h), t21((size_t)w * h);
    fh_transform(d12.data(), w, w, h, norm, t12.data(), w);
    fh_transform(d21.data(), w, w, h, norm, t21.data(), w);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            mask[(long)y * mpitch + x] = t12[(size_t)y * w + x] < t21[(size_t)y * w + x] ? 255 : 0;
    // excessive overlap: fewer exclusive pixels than the threshold in both
    const long threshold = 2L * 2L * ((long)h + (long)w);
    return std::max(n12, n21) < threshold ? -2 : 0;
}

// nearestFeatureTransform with strip boundaries (src/nearest.h:474-520):
// each difference image copied twice side by side (horizontal strip, 1), one
// above the other (vertical strip, 2) or four times (double strip, 3)
// (quadruple_image, src/nearest.h:138-185), transformed with open
// boundaries, and cut back to w x h (quater_image, src/nearest.h:200-285):
// a doubled axis of length n takes position p + n for p < n / 2 and p
// otherwise.  The overlap tally is that of the original difference images.
int orc_nft_mask_periodic(const uint8_t* white, long wpitch, const uint8_t* black, long bpitch, int w, int h,
                          int norm, int boundary, uint8_t* mask, long mpitch)
{
    if (w < 1 || h < 1 || (norm != 1 && norm != 2) || boundary < 1 || boundary > 3) return -1;
    const bool hx = boundary == 1 || boundary == 3, hy = boundary == 2 || boundary == 3;
    const int W = hx ? 2 * w : w, H = hy ? 2 * h : h;
    std::vector<uint8_t> d12((size_t)W * H), d21((size_t)W * H);
    long n12 = 0, n21 = 0;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const uint8_t a = white[(long)y * wpitch + x], b = black[(long)y * bpitch + x];
            const uint8_t v12 = b < a ? (uint8_t)(a - b) : 0, v21 = a < b ? (uint8_t)(b - a) : 0;
            n12 += v12 != 0;
            n21 += v21 != 0;
            for (int cy = 0; cy < (hy ? 2 : 1); ++cy)
                for (int cx = 0; cx < (hx ? 2 : 1); ++cx) {
                    const size_t i = (size_t)(y + cy * h) * W + (x + cx * w);
                    d12[i] = v12;
                    d21[i] = v21;
                }
        }
    std::vector<float> t12((size_t)W * H), t21((size_t)W * H);
    fh_transform(d12.data(), W, W, H, norm, t12.data(), W);
    fh_transform(d21.data(), W, W, H, norm, t21.data(), W);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const int X = hx && x < w / 2 ? x + w : x, Y = hy && y < h / 2 ? y + h : y;
            const size_t i = (size_t)Y * W + X;
            mask[(long)y * mpitch + x] = t12[i] < t21[i] ? 255 : 0;
        }
    const long threshold = 2L * 2L * ((long)h + (long)w);
    return std::max(n12, n21) < threshold ? -2 : 0;
}

}  // extern "C"
