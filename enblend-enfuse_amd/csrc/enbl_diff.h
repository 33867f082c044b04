// Kernels of the relaxed differentiable enfuse (diff.hip): forward terms and
// hand-written adjoints, all in double.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace enbl {

constexpr int kDiffMaxK = 16;   // images per differentiable fusion

struct DPlane {                 // one double plane
    const double* p;
    long pitch;
};

// Planes of one pyramid level for every image (level l and l+1 of the image
// Gaussian pyramids, level l of the mask pyramids).
struct DiffLevel {
    int K;
    int fw, fh, cw, ch;                     // level l and l+1 sizes (cw = 0: coarsest)
    long gpitch, gplane, cpitch, cplane;    // image pyramid level l / l+1 geometry
    long mpitch;                            // mask level l pitch
    const double* g[kDiffMaxK];             // G_k level l (3 planes)
    const double* gc[kDiffMaxK];            // G_k level l+1
    const double* m[kDiffMaxK];             // M_k level l
};

// Per-image weight terms of the forward pass (planes of pitch tp).
struct DiffTerms {
    int K, w, h;
    long tp;
    double* gray[kDiffMaxK];
    double* E[kDiffMaxK];
    double* s[kDiffMaxK];
    double* sd[kDiffMaxK];
    double* mean[kDiffMaxK];
};

int launch_diff_terms(const double* img, int w, int h, double mu, double sigma, double* gray, double* E, double* s,
                      long tp, hipStream_t st);
int launch_diff_sd(const double* gray, int w, int h, long tp, int n, double* sd, double* mean, hipStream_t st);
int launch_diff_weights(const DiffTerms& t, double we, double ws, double wc, double* const* mhat, long mpitch,
                        double* S, hipStream_t st);
int launch_diff_log(const double* img, int w, int h, double* g, long pitch, long plane, hipStream_t st);
int launch_diff_merge(const DiffLevel& a, double* r, long rpitch, long rplane, hipStream_t st);
int launch_diff_out(const double* c, long pitch, long plane, int w, int h, double* out, hipStream_t st);
// adjoint
int launch_diff_gout(const double* gout, const double* c, long pitch, long plane, int w, int h, double* gc,
                     hipStream_t st);
int launch_diff_expand_T(const double* g, long gpitch, long gplane, int fw, int fh, int C, double* gc, long cpitch,
                         long cplane, int cw, int ch, hipStream_t st);
int launch_diff_reduce_T(const double* g, long gpitch, long gplane, int dw, int dh, int C, double* gf, long fpitch,
                         long fplane, int sw, int sh, hipStream_t st);

// One level l of one image's pyramid adjoint:
//   gG_l = gC_l M_l - expand^T(gC_l-1 M_l-1) + reduce^T(gG_l+1)
//   gM_l = sum_c gC_l (G_l - expand(G_l+1)) + reduce^T(gM_l+1)
// At level 0, the image gradient gI = gG_0 dT(I) is written instead of gG_0.
struct DiffAdjLevel {
    int l, L;
    int fw, fh, cw, ch, pw, ph;             // level l, l+1, l-1 sizes
    const double* G; long gpitch, gplane;   // G_l
    const double* Gc; long gcpitch, gcplane;   // G_l+1
    const double* M; long mpitch;           // M_l
    const double* Mp; long mppitch;         // M_l-1
    const double* gC; long gcp, gcpl;       // gC_l
    const double* gCp; long gcpp, gcppl;    // gC_l-1
    const double* gGc; long ggcp, ggcpl;    // gG_l+1 (total)
    const double* gMc; long gmcp;           // gM_l+1 (total)
    double* gG; long ggp, ggpl;             // out: gG_l (l > 0)
    double* gM; long gmp;                   // out: gM_l (total)
    const double* img; double* gimg;        // level 0: the image and its gradient (h, w, 3)
};
int launch_diff_adj_level(const DiffAdjLevel& a, hipStream_t st);

// Normalisation + weight adjoint for every image: parameter partials into
// gparam[5] (atomic), pointwise image gradients added to gimgs, and the
// contrast coefficient A = gm w_c / ((N - 1) sd) of every window centre (0 off
// the interior) for the window pass.
struct DiffWeightAdj {
    int K, w, h, n;
    long tp;
    double we, ws, wc, mu, sigma;
    const double* img[kDiffMaxK];           // (h, w, 3)
    double* gimg[kDiffMaxK];
    const double* gmhat[kDiffMaxK]; long gmpitch;   // total mask gradient at level 0
    const double* mhat[kDiffMaxK]; long mpitch;
    const double* S;
    const double* gray[kDiffMaxK];
    const double* E[kDiffMaxK];
    const double* s[kDiffMaxK];
    const double* sd[kDiffMaxK];
    const double* mean[kDiffMaxK];
    double* A[kDiffMaxK];
    double* gparam;
};
int launch_diff_weight_adj(const DiffWeightAdj& a, hipStream_t st);
// gimg += (sum over the windows p containing q of A_p (gray_q - mean_p)) / 3.
int launch_diff_contrast_adj(const double* A, const double* mean, const double* gray, int w, int h, long tp, int n,
                             double* gimg, hipStream_t st);

}  // namespace enbl
