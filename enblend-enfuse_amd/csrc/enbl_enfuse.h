// Launchers of the enfuse weighting and merge kernels (enfuse.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/enblend_amd.h"

namespace enbl {

// A grayscale projector: ENBL_GRAY_* and the normalised mixer weights.
struct GrayProj {
    int kind;
    int pad_;
    double wr, wg, wb;
};

struct WeightArgs {
    double w_exposure, w_saturation, w_contrast, w_entropy;
    double exposure_optimum, exposure_width;   // width already rescaled to the function's FWHM
    int exposure_function;
    int exposure_cutoff;                       // CutoffExposureFunctor: min channel >= lo, max channel <= hi
    double exposure_lo, exposure_hi;
    GrayProj proj;                             // --gray-projector of the exposure gray
};

int launch_gray(int itype, const void* img, long ip, int w, int h, void* gray, long gp, const GrayProj& pj,
                hipStream_t s);
// localStdDevIf of K gray planes (plane m at gray + m * gstride elements,
// result plane m at out + m * ostride), each under its own alpha.
int launch_stddev(int itype, const void* gray, long gp, size_t gstride, const uint8_t* const* alphas,
                  const long* apitches, int K, int w, int h, int n, void* out, long op, size_t ostride,
                  hipStream_t s);
int launch_entropy(int itype, const void* img, long ip, const uint8_t* alpha, long ap, int w, int h, int n,
                   const double* tab, void* out, long op, hipStream_t s);
// The local entropy (windows 3, 5 and 7) and the whole weights pass in one
// launch: mask, norm and union alpha as launch_weights writes them, no
// entropy plane.
int launch_entropy_weights(int itype, const void* img, long ip, const uint8_t* alpha, long ap, int w, int h, int n,
                           const double* tab, const WeightArgs& a, const float* lut, const void* grad, long gp,
                           float* mask, long mp, float* norm, long np_, uint8_t* ualpha, long up, hipStream_t s);
int launch_weights(int itype, const void* img, long ip, const uint8_t* alpha, long ap, int w, int h,
                   const WeightArgs& a, const float* lut, const void* grad, long gp, const void* ent, long ep,
                   float* mask, long mp, float* norm, long np_, uint8_t* ualpha, long up, hipStream_t s);
int launch_normalize(int mtype, const float* mask, long mp, const float* norm, long np_, int w, int h,
                     float fallback, double scale, int hard, void* out, long op, hipStream_t s);
// Hard masks over K weight planes (plane m at masks + m * plane).
int launch_hard_mask(float* masks, long mp, size_t plane, int w, int h, int K, hipStream_t s);
int launch_fill(void* p, size_t bytes, hipStream_t s);

// Merges of n <= kMergeMax images' Laplacian levels (3 planes each, one
// geometry) weighted by their mask levels into the result level, in image
// order; `first`: the result level is assigned by the first image.
constexpr int kMergeMax = 8;
struct MergeMany {
    const void* lp[kMergeMax];
    const void* g[kMergeMax];
    long lpp, lpl, gpp;
    void* res;
    long rp, rl;
    double maxpyr;
    int w, h, n, first;
};
int launch_merge_many(int ptype, int mtype, const MergeMany& a, hipStream_t s);
// The image clamped for --entropy-cutoff (src/enfuse.h:1179-1185), packed
// (pitch 3 w): channel <= lo -> 0 (the blue channel tests red), >= up -> max.
int launch_entropy_clamp(int itype, const void* img, long ip, int w, int h, double lo, double up, void* out,
                         hipStream_t s);

// Contrast by Laplacian-of-Gaussian edges (src/enfuse.h:1020-1093): the
// separable passes of VIGRA's laplacianOfGaussian / gaussianSharpening with
// mirrored borders.  Planes of the gray type G (int32 for integer images,
// float for float images) and of its real promote R (double, float).
struct EdgePlanes {
    const void* src; long spitch;          // G: the gray (or sharpened gray) plane
    void* tA; void* tB; long tpitch;       // R scratch planes
    int w, h;
};
// Row pass: tA = src (*) ka along x (and tB = src (*) kb when kb != NULL).
int launch_edge_rows(int itype, const EdgePlanes& e, const double* ka, int ra, const double* kb, int rb,
                     hipStream_t s);
// Column pass of gaussianSharpening: out = cast_G((1 + f) gray - f (tA (*) k along y)).
int launch_edge_sharpen(int itype, const EdgePlanes& e, const double* k, int r, double f, const void* gray,
                        long gpitch, void* out, long opitch, hipStream_t s);
// Column passes of laplacianOfGaussian on tA = src (*) d2 and tB = src (*)
// smooth (rows), the magnitude and the clamping (min_curvature <= 0) or
// fill-in functor (lc: the local standard deviation), written where alpha
// is set, 0 elsewhere.
int launch_edge_log(int itype, const EdgePlanes& e, const double* ks, int rs, const double* kd, int rd,
                    const uint8_t* alpha, long apitch, const void* lc, long lpitch, double min_curvature,
                    double scalar_max, void* grad, long gpitch, hipStream_t s);

namespace host {
// Parameter checks of enfuse (the reference's preconditions and option checks).
int check_params(int itype, int w, int h, const enbl_enfuse_params* p);
WeightArgs weight_args(const enbl_enfuse_params* p);
// (float)(W_exposure * weight(code / max)) for every gray code of an integer image type.
void exposure_lut(int itype, const enbl_enfuse_params* p, const WeightArgs& a, std::vector<float>& lut);
}  // namespace host

}  // namespace enbl
