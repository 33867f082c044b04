// C ABI of the relaxed differentiable enfuse (declared in
// include/enblend_amd.h): forward, adjoint, and the transposed pyramid
// operators.  The backward call recomputes the forward's Gaussian pyramids and
// weight terms and evaluates every Laplacian level on the fly, so no state
// passes between the two calls.
#include <math.h>

#include <vector>

#include "enbl_diff.h"
#include "enbl_enfuse.h"
#include "enbl_internal.h"

using namespace enbl;
using namespace enbl::host;

namespace {

struct Fwd {
    int w, h, K, L, n;
    long tp;
    double* terms = nullptr;     // 5 planes per image: gray, E, s, sd, mean
    double* S = nullptr;
    uint8_t* ones = nullptr;
    enbl_pyramid* G[kDiffMaxK];
    enbl_pyramid* M[kDiffMaxK];
    enbl_pyramid* C = nullptr;
    double* plane(int k, int which) const { return terms + ((size_t)k * 5 + which) * (size_t)tp * h; }
};

int check(int w, int h, int K, int levels, int n, const enbl_diff_params* p)
{
    if (!p) return fail(ENBL_E_ARG, "enbl_diff_enfuse: null parameters");
    if (w < 1 || h < 1) return fail(ENBL_E_SIZE, "enbl_diff_enfuse: empty image");
    if (K < 1 || K > kDiffMaxK) return fail(ENBL_E_UNSUPPORTED, "enbl_diff_enfuse: %d images (1..%d)", K, kDiffMaxK);
    if (levels < 1 || levels > 29) return fail(ENBL_E_PRECONDITION, "filterHalfWidth: levels outside of permissible range");
    if (n < 3 || !(n & 1)) return fail(ENBL_E_ARG, "enbl_diff_enfuse: contrast window must be odd and >= 3");
    if (!(p->exposure_width > 0.0)) return fail(ENBL_E_ARG, "width_parameter");
    int lw = w, lh = h;
    for (int l = 0; l + 1 < levels; ++l) {
        if (lw < 2 || lh < 2) return fail(ENBL_E_PRECONDITION, "src image too small in reduce");
        lw = (lw + 1) >> 1;
        lh = (lh + 1) >> 1;
    }
    return ENBL_OK;
}

double* lvd(const enbl_pyramid* p, int l) { return static_cast<double*>(level_ptr(p, l)); }

int forward(enbl_context* ctx, int w, int h, int K, const double* imgs, const enbl_diff_params* p, int L, int n, Fwd& F)
{
    const hipStream_t s = ctx->stream;
    F.w = w; F.h = h; F.K = K; F.L = L; F.n = n;
    F.tp = (long)round_up((size_t)w, 64);
    const size_t pl = (size_t)F.tp * h;
    void *terms, *Sbuf;
    int rc = scratch(ctx, 9, pl * 8 * 5 * (size_t)K, &terms);
    if (!rc) rc = scratch(ctx, 10, pl * 8 + pl, &Sbuf);
    if (rc) return rc;
    F.terms = (double*)terms;
    F.S = (double*)Sbuf;
    F.ones = (uint8_t*)Sbuf + pl * 8;
    rc = hip_check(hipMemsetAsync(F.ones, 255, pl, s), "alpha plane");
    if (rc) return rc;
    for (int k = 0; k < K; ++k) {
        rc = cached_pyramid(ctx, kRoleImage, ENBL_F64, 3, w, h, L, k, &F.G[k]);
        if (!rc) rc = cached_pyramid(ctx, kRoleMask, ENBL_F64, 1, w, h, L, k, &F.M[k]);
        if (rc) return rc;
    }
    rc = cached_pyramid(ctx, kRoleResult, ENBL_F64, 3, w, h, L, 0, &F.C);
    if (rc) return rc;
    const size_t ipx = (size_t)w * h * 3;
    DiffTerms t{};
    t.K = K; t.w = w; t.h = h; t.tp = F.tp;
    {
        Timed tm(ctx, "diff_terms", (double)K * w * h * (24.0 + 5 * 8.0) + (double)K * w * h * 8.0 * (n * n + 2));
        for (int k = 0; k < K; ++k) {
            t.gray[k] = F.plane(k, 0); t.E[k] = F.plane(k, 1); t.s[k] = F.plane(k, 2);
            t.sd[k] = F.plane(k, 3); t.mean[k] = F.plane(k, 4);
            launch_diff_terms(imgs + k * ipx, w, h, p->exposure_optimum, p->exposure_width, t.gray[k], t.E[k], t.s[k], F.tp, s);
            launch_diff_sd(t.gray[k], w, h, F.tp, n, t.sd[k], t.mean[k], s);
        }
        rc = launch_status("diff terms");
        if (rc) return rc;
    }
    {
        Timed tm(ctx, "diff_weights", (double)K * w * h * (24.0 + 8.0) + w * h * 8.0);
        double* mh[kDiffMaxK];
        for (int k = 0; k < K; ++k) mh[k] = lvd(F.M[k], 0);
        launch_diff_weights(t, p->w_exposure, p->w_saturation, p->w_contrast, mh, F.M[0]->lv[0].pitch, F.S, s);
        rc = launch_status("diff weights");
        if (rc) return rc;
    }
    for (int k = 0; k < K; ++k) {
        const Level& L0 = F.G[k]->lv[0];
        {
            Timed tm(ctx, "diff_log", (double)w * h * 48.0);
            launch_diff_log(imgs + k * ipx, w, h, lvd(F.G[k], 0), L0.pitch, L0.plane, s);
            rc = launch_status("diff log");
            if (rc) return rc;
        }
        rc = gaussian_levels(ctx, F.G[k], 0, F.ones, F.tp);
        if (!rc) rc = gaussian_levels(ctx, F.M[k], 0, F.ones, F.tp);
        if (rc) return rc;
    }
    for (int l = 0; l < L; ++l) {
        const Level& a = F.G[0]->lv[l];
        DiffLevel d{};
        d.K = K;
        d.fw = a.w; d.fh = a.h;
        d.cw = l + 1 < L ? F.G[0]->lv[l + 1].w : 0;
        d.ch = l + 1 < L ? F.G[0]->lv[l + 1].h : 0;
        d.gpitch = a.pitch; d.gplane = a.plane;
        if (l + 1 < L) { d.cpitch = F.G[0]->lv[l + 1].pitch; d.cplane = F.G[0]->lv[l + 1].plane; }
        d.mpitch = F.M[0]->lv[l].pitch;
        for (int k = 0; k < K; ++k) {
            d.g[k] = lvd(F.G[k], l);
            d.gc[k] = l + 1 < L ? lvd(F.G[k], l + 1) : nullptr;
            d.m[k] = lvd(F.M[k], l);
        }
        Timed tm(ctx, "diff_merge", (double)a.w * a.h * (K * (24.0 + 8.0) + 24.0) + (l + 1 < L ? (double)K * d.cw * d.ch * 24.0 : 0.0));
        launch_diff_merge(d, lvd(F.C, l), F.C->lv[l].pitch, F.C->lv[l].plane, s);
        rc = launch_status("diff merge");
        if (rc) return rc;
    }
    return enbl_collapse_pyramid(ctx, 0, F.C);
}

}  // namespace

extern "C" {

int enbl_diff_enfuse_forward(enbl_context* ctx, int w, int h, int K, const double* imgs, const enbl_diff_params* p,
                             int levels, int n, double* out)
{
    if (!ctx || !imgs || !out) return fail(ENBL_E_ARG, "enbl_diff_enfuse_forward: null argument");
    int rc = check(w, h, K, levels, n, p);
    if (rc) return rc;
    Fwd F;
    rc = forward(ctx, w, h, K, imgs, p, levels, n, F);
    if (rc) return rc;
    Timed tm(ctx, "diff_out", (double)w * h * 48.0);
    launch_diff_out(lvd(F.C, 0), F.C->lv[0].pitch, F.C->lv[0].plane, w, h, out, ctx->stream);
    return launch_status("diff out");
}

int enbl_diff_enfuse_backward(enbl_context* ctx, int w, int h, int K, const double* imgs, const enbl_diff_params* p,
                              int levels, int n, const double* gout, double* gimgs, double* gparams)
{
    if (!ctx || !imgs || !gout || !gimgs || !gparams) return fail(ENBL_E_ARG, "enbl_diff_enfuse_backward: null argument");
    int rc = check(w, h, K, levels, n, p);
    if (rc) return rc;
    Fwd F;
    rc = forward(ctx, w, h, K, imgs, p, levels, n, F);
    if (rc) return rc;
    const hipStream_t s = ctx->stream;
    const int L = levels;
    const size_t ipx = (size_t)w * h * 3;
    // Output transform and collapse adjoint: gC_l, finest first.
    enbl_pyramid* gC = nullptr;
    rc = cached_pyramid(ctx, kRoleAdjoint, ENBL_F64, 3, w, h, L, 0, &gC);
    if (rc) return rc;
    {
        Timed tm(ctx, "diff_gout", (double)w * h * 72.0);
        launch_diff_gout(gout, lvd(F.C, 0), F.C->lv[0].pitch, F.C->lv[0].plane, w, h, lvd(gC, 0), s);
        rc = launch_status("diff gout");
        if (rc) return rc;
    }
    for (int l = 0; l + 1 < L; ++l) {
        const Level& a = gC->lv[l];
        const Level& b = gC->lv[l + 1];
        rc = launch_fill(lvd(gC, l + 1), (size_t)b.plane * 3 * 8, s);
        if (rc) return rc;
        Timed tm(ctx, "diff_expand_T", (double)a.w * a.h * 24.0 + (double)b.w * b.h * 48.0);
        launch_diff_expand_T(lvd(gC, l), a.pitch, a.plane, a.w, a.h, 3, lvd(gC, l + 1), b.pitch, b.plane, b.w, b.h, s);
        rc = launch_status("diff expand_T");
        if (rc) return rc;
    }
    // Per image: merge, Laplacian, Gaussian and mask pyramid adjoints.
    enbl_pyramid* gM[kDiffMaxK];
    for (int k = 0; k < K; ++k) {
        rc = cached_pyramid(ctx, kRoleAdjoint, ENBL_F64, 1, w, h, L, k, &gM[k]);
        if (rc) return rc;
    }
    void* gGbuf = nullptr;
    const size_t gGsz = L > 1 ? (size_t)F.G[0]->lv[1].plane * 3 : 1;
    rc = scratch(ctx, 12, 2 * gGsz * 8, &gGbuf);
    if (rc) return rc;
    double* gG[2] = {(double*)gGbuf, (double*)gGbuf + gGsz};
    for (int k = 0; k < K; ++k) {
        for (int l = L - 1; l >= 0; --l) {
            const Level& a = F.G[k]->lv[l];
            DiffAdjLevel d{};
            d.l = l; d.L = L;
            d.fw = a.w; d.fh = a.h;
            d.G = lvd(F.G[k], l); d.gpitch = a.pitch; d.gplane = a.plane;
            d.M = lvd(F.M[k], l); d.mpitch = F.M[k]->lv[l].pitch;
            d.gC = lvd(gC, l); d.gcp = gC->lv[l].pitch; d.gcpl = gC->lv[l].plane;
            if (l + 1 < L) {
                const Level& b = F.G[k]->lv[l + 1];
                d.cw = b.w; d.ch = b.h;
                d.Gc = lvd(F.G[k], l + 1); d.gcpitch = b.pitch; d.gcplane = b.plane;
                d.gGc = gG[(l + 1) & 1]; d.ggcp = b.pitch; d.ggcpl = b.plane;
                d.gMc = lvd(gM[k], l + 1); d.gmcp = gM[k]->lv[l + 1].pitch;
            }
            if (l >= 1) {
                const Level& f = F.G[k]->lv[l - 1];
                d.pw = f.w; d.ph = f.h;
                d.Mp = lvd(F.M[k], l - 1); d.mppitch = F.M[k]->lv[l - 1].pitch;
                d.gCp = lvd(gC, l - 1); d.gcpp = gC->lv[l - 1].pitch; d.gcppl = gC->lv[l - 1].plane;
                d.gG = gG[l & 1]; d.ggp = a.pitch; d.ggpl = a.plane;
            } else {
                d.img = imgs + k * ipx;
                d.gimg = gimgs + k * ipx;
            }
            d.gM = lvd(gM[k], l); d.gmp = gM[k]->lv[l].pitch;
            Timed tm(ctx, "diff_adj_level", (double)a.w * a.h * (24.0 * 4 + 8.0 * 3) + (l + 1 < L ? (double)d.cw * d.ch * 56.0 : 0.0));
            launch_diff_adj_level(d, s);
            rc = launch_status("diff adjoint level");
            if (rc) return rc;
        }
    }
    // Normalisation and weight adjoint.
    void *Abuf, *gpbuf;
    const size_t pl = (size_t)F.tp * h;
    rc = scratch(ctx, 11, pl * 8 * (size_t)K, &Abuf);
    if (!rc) rc = scratch(ctx, 13, 64, &gpbuf);
    if (rc) return rc;
    rc = launch_fill(gpbuf, 64, s);
    if (rc) return rc;
    DiffWeightAdj wa{};
    wa.K = K; wa.w = w; wa.h = h; wa.n = n; wa.tp = F.tp;
    wa.we = p->w_exposure; wa.ws = p->w_saturation; wa.wc = p->w_contrast;
    wa.mu = p->exposure_optimum; wa.sigma = p->exposure_width;
    wa.gmpitch = gM[0]->lv[0].pitch;
    wa.mpitch = F.M[0]->lv[0].pitch;
    wa.S = F.S;
    wa.gparam = (double*)gpbuf;
    for (int k = 0; k < K; ++k) {
        wa.img[k] = imgs + k * ipx;
        wa.gimg[k] = gimgs + k * ipx;
        wa.gmhat[k] = lvd(gM[k], 0);
        wa.mhat[k] = lvd(F.M[k], 0);
        wa.gray[k] = F.plane(k, 0); wa.E[k] = F.plane(k, 1); wa.s[k] = F.plane(k, 2);
        wa.sd[k] = F.plane(k, 3); wa.mean[k] = F.plane(k, 4);
        wa.A[k] = (double*)Abuf + (size_t)k * pl;
    }
    {
        Timed tm(ctx, "diff_weight_adj", (double)K * w * h * (8.0 * 7 + 24.0 * 3));
        launch_diff_weight_adj(wa, s);
        rc = launch_status("diff weight adjoint");
        if (rc) return rc;
    }
    for (int k = 0; k < K; ++k) {
        Timed tm(ctx, "diff_contrast_adj", (double)w * h * (8.0 * (2 * n * n + 1) + 48.0));
        launch_diff_contrast_adj(wa.A[k], wa.mean[k], wa.gray[k], w, h, F.tp, n, wa.gimg[k], s);
        rc = launch_status("diff contrast adjoint");
        if (rc) return rc;
    }
    return hip_check(hipMemcpyAsync(gparams, gpbuf, 5 * sizeof(double), hipMemcpyDeviceToDevice, s), "parameter gradient");
}

int enbl_expand_adjoint(enbl_context* ctx, const double* g, long gpitch, int fw, int fh, double* gc, long cpitch, int cw,
                        int ch)
{
    if (!ctx || !g || !gc) return fail(ENBL_E_ARG, "enbl_expand_adjoint: null argument");
    if (cw != (fw + 1) / 2 || ch != (fh + 1) / 2 || cw < 1 || ch < 1) return fail(ENBL_E_SIZE, "enbl_expand_adjoint: sizes");
    launch_diff_expand_T(g, gpitch, 0, fw, fh, 1, gc, cpitch, 0, cw, ch, ctx->stream);
    return launch_status("expand adjoint");
}

int enbl_reduce_adjoint(enbl_context* ctx, const double* g, long gpitch, int dw, int dh, double* gf, long fpitch, int sw,
                        int sh)
{
    if (!ctx || !g || !gf) return fail(ENBL_E_ARG, "enbl_reduce_adjoint: null argument");
    if (dw != (sw + 1) / 2 || dh != (sh + 1) / 2 || sw < 2 || sh < 2) return fail(ENBL_E_SIZE, "enbl_reduce_adjoint: sizes");
    launch_diff_reduce_T(g, gpitch, 0, dw, dh, 1, gf, fpitch, 0, sw, sh, ctx->stream);
    return launch_status("reduce adjoint");
}

}  // extern "C"
