// C ABI of enfuse's weighting and fusion (declared in include/enblend_amd.h).
//
// enbl_enfuse follows enfuseMain (src/enfuse.h:1217-1609, soft or hard masks) stage by stage but
// keeps every intermediate on the device: the K weight images and the norm
// image are built first (the reference's first pass over the inputs), then
// each input's Laplacian pyramid and normalised mask Gaussian pyramid are
// merged level by level into the result pyramid, which is collapsed and
// converted once.
#include <limits>
#include <float.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "enbl_batch.h"
#include "enbl_enfuse.h"
#include "enbl_internal.h"

using namespace enbl;
using namespace enbl::host;

namespace enbl {
namespace host {

const double kFwhmGaussian = 2.3548200450309493820231386529193992755;
const double kFwhm[] = {kFwhmGaussian, 2.8284271247461900976033774484193961571,
                        2.0943951023931954923084289221863352561, M_PI,
                        1.6817928305074290860622509524664297901};

// Built-in ExposureWeight::weight (src/exposure_weight.h:44-120), evaluated
// on the host for the integer-image tables.
double exposure_host(int fn, double opt, double width, double y)
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

// VIGRA's Kernel1D<double>::initGaussian (order 0) and
// initGaussianDerivative (order 2) with the default window: radius
// (int)(3 sigma + 0.5), resp. (int)(4 sigma + 0.5), at least 1; samples of
// the Gaussian (or its second derivative, (1 - (x/sigma)^2) g) from -radius
// up, the derivative's mean removed, then scaled so that sum k = 1, resp.
// sum k x^2 / 2 = 1.
int vigra_radius(double sigma, int order)
{
    const int r = order == 0 ? (int)(3.0 * sigma + 0.5) : (int)((3.0 + 0.5 * order) * sigma + 0.5);
    return r == 0 ? 1 : r;
}

std::vector<double> vigra_kernel(double sigma, int order)
{
    const int radius = vigra_radius(sigma, order);
    const double sigma2 = -0.5 / sigma / sigma;
    const double norm = order == 0 ? 1.0 / (sqrt(2.0 * M_PI) * sigma)
                                   : -1.0 / (sqrt(2.0 * M_PI) * (sigma * sigma) * sigma);
    std::vector<double> k;
    double dc = 0.0;
    for (double x = -(double)radius; x <= (double)radius; ++x) {
        const double g = norm * exp(x * x * sigma2);
        const double v = order == 0 ? g : (1.0 - (x / sigma) * (x / sigma)) * g;
        k.push_back(v);
        dc += v;
    }
    double sum = 0.0;
    if (order == 0) {
        for (double v : k) sum += v;
    } else {
        dc = dc / (2.0 * radius + 1.0);
        for (double& v : k) v -= dc;
        double x = -(double)radius;
        for (size_t i = 0; i < k.size(); ++i, ++x) sum = sum + k[i] * ((-x) * (-x)) / 2.0;
    }
    sum = 1.0 / sum;
    for (double& v : k) v = v * sum;
    return k;
}

GrayProj gray_proj(const enbl_enfuse_params* p)
{
    GrayProj g;
    memset(&g, 0, sizeof g);
    g.kind = p->gray_projector;
    if (g.kind == ENBL_GRAY_MIXER) {
        const double sum = p->mixer_red + p->mixer_green + p->mixer_blue;
        g.wr = p->mixer_red / sum;
        g.wg = p->mixer_green / sum;
        g.wb = p->mixer_blue / sum;
    }
    return g;
}

int check_params(int itype, int w, int h, const enbl_enfuse_params* p)
{
    if (!p) return fail(ENBL_E_ARG, "enfuse: null parameters");
    if (itype != ENBL_U8 && itype != ENBL_U16 && itype != ENBL_F32) return fail(ENBL_E_UNSUPPORTED, "enfuse: image type %d", itype);
    if (p->hard_mask != 0 && p->hard_mask != 1) return fail(ENBL_E_ARG, "enfuse: hard_mask %d", p->hard_mask);
    if (p->gray_projector < ENBL_GRAY_AVERAGE || p->gray_projector > ENBL_GRAY_MIXER)
        return fail(ENBL_E_ARG, "unknown grayscale projector %d", p->gray_projector);
    if (p->gray_projector == ENBL_GRAY_MIXER &&
        (p->mixer_red < 0.0 || p->mixer_green < 0.0 || p->mixer_blue < 0.0 ||
         p->mixer_red + p->mixer_green + p->mixer_blue == 0.0))
        return fail(ENBL_E_ARG, "channel-mixer: negative weight or zero sum of weights");
    if (p->exposure_cutoff && p->w_exposure > 0.0) {
        if (p->exposure_lower_cutoff < 0.0) return fail(ENBL_E_PRECONDITION, "negative lower exposure cutoff");
        if (p->exposure_upper_cutoff < 0.0) return fail(ENBL_E_PRECONDITION, "negative upper exposure cutoff");
        if (p->exposure_lower_cutoff > p->exposure_upper_cutoff)
            return fail(ENBL_E_PRECONDITION, "lower exposure cutoff exceeds upper cutoff");
    }
    if (p->entropy_cutoff && p->w_entropy > 0.0) {
        if (p->entropy_lower_cutoff < 0.0) return fail(ENBL_E_PRECONDITION, "negative lower entropy cutoff");
        if (p->entropy_upper_cutoff < 0.0) return fail(ENBL_E_PRECONDITION, "negative upper entropy cutoff");
        if (p->entropy_lower_cutoff > p->entropy_upper_cutoff)
            return fail(ENBL_E_PRECONDITION, "lower entropy cutoff exceeds upper cutoff");
    }
    if (p->exposure_function < 0 || p->exposure_function > ENBL_EXPOSURE_BISQUARE)
        return fail(ENBL_E_ARG, "enfuse: exposure function %d", p->exposure_function);
    if (p->w_exposure > 0.0 && (p->exposure_optimum < 0.0 || p->exposure_optimum > 1.0))
        return fail(ENBL_E_ARG, "y_optimum");       // ExposureWeight::check_invariant
    if (p->w_exposure > 0.0 && !(p->exposure_width > 0.0))
        return fail(ENBL_E_ARG, "width_parameter");
    if (p->w_exposure > 0.0 && p->exposure_table) {
        // A user weight function reaches the device only as its table over
        // the gray codes of an integer image; a float image would need
        // weight(gray / FLT_MAX) per pixel (src/enfuse.h:637-640).
        if (itype == ENBL_F32)
            return fail(ENBL_E_UNSUPPORTED, "exposure weight table given for a float image");
        // check_weight_function (src/exposure_weight.cc:237-264) over the
        // table's points; enfuse exits on any result but OK
        // (src/enfuse.cc:1920-1945).
        const int n = itype == ENBL_U8 ? 256 : 65536;
        int zeros = 0;
        for (int g = 0; g < n; ++g) {
            const double wt = p->exposure_table[g];
            if (wt < 0.0) return fail(ENBL_E_PRECONDITION, "unusable exposure weight function: at least one weight is negative");
            if (wt > 1.0) return fail(ENBL_E_PRECONDITION, "unusable exposure weight function: at least one weight is larger than one");
            if (wt >= 0.0 && wt <= std::numeric_limits<double>::denorm_min()) ++zeros;
        }
        if (zeros == n) return fail(ENBL_E_PRECONDITION, "unusable exposure weight function: too many zeros");
    }
    if (p->w_contrast > 0.0 && !(p->edge_scale >= 0.0)) return fail(ENBL_E_ARG, "enfuse: edge scale");
    if (p->w_contrast > 0.0 && (p->edge_scale <= 0.0 || p->min_curvature > 0.0)) {
        const int n = p->contrast_window;
        if (n < 3 || !(n & 1)) return fail(ENBL_E_ARG, "enfuse: contrast window must be odd and >= 3");
        if (w < n || h < n) return fail(ENBL_E_PRECONDITION, "localStdDevIf(): window larger than image");
    }
    if (p->w_contrast > 0.0 && p->edge_scale > 0.0) {
        // kernels of more taps than any line could hold are rejected before
        // their radius is formed
        if (p->edge_scale > 1e6 || p->lce_scale > 1e6)
            return fail(ENBL_E_PRECONDITION, "convolveLine(): kernel longer than line");
        if (!(p->lce_scale >= 0.0) || !(p->lce_factor >= 0.0))
            return fail(ENBL_E_PRECONDITION, "gaussianSharpening(): scale and sharpening factor must be >= 0");
        const int rs = vigra_radius(p->edge_scale, 0), rd = vigra_radius(p->edge_scale, 2);
        const int rl = p->lce_scale > 0.0 ? vigra_radius(p->lce_scale, 0) : 0;
        if (std::max(std::max(rs, rd), rl) >= std::min(w, h))
            return fail(ENBL_E_PRECONDITION, "convolveLine(): kernel longer than line");
    }
    if (p->w_entropy > 0.0) {
        const int n = p->entropy_window;
        if (n < 3 || !(n & 1) || n * n > 121) return fail(ENBL_E_UNSUPPORTED, "enfuse: entropy window %d", n);
        if (w < n || h < n) return fail(ENBL_E_PRECONDITION, "localEntropyIf(): window larger than image");
    }
    return ENBL_OK;
}

WeightArgs weight_args(const enbl_enfuse_params* p)
{
    WeightArgs a;
    memset(&a, 0, sizeof a);
    a.w_exposure = p->w_exposure;
    a.w_saturation = p->w_saturation;
    a.w_contrast = p->w_contrast;
    a.w_entropy = p->w_entropy;
    a.exposure_optimum = p->exposure_optimum;
    a.exposure_function = p->exposure_function;
    a.exposure_width = p->exposure_function == 0 ? p->exposure_width
                                                 : p->exposure_width * kFwhmGaussian / kFwhm[p->exposure_function];
    a.exposure_cutoff = p->exposure_cutoff ? 1 : 0;
    a.exposure_lo = p->exposure_cutoff ? p->exposure_lower_cutoff : 0.0;
    a.exposure_hi = p->exposure_cutoff ? p->exposure_upper_cutoff : 0.0;
    a.proj = gray_proj(p);
    return a;
}

void exposure_lut(int itype, const enbl_enfuse_params* p, const WeightArgs& a, std::vector<float>& lut)
{
    const int n = itype == ENBL_U8 ? 256 : 65536;
    const double maxv = itype == ENBL_U8 ? 255.0 : 65535.0;
    lut.resize((size_t)n);
    for (int g = 0; g < n; ++g) {
        const double wt = p->exposure_table ? p->exposure_table[g]
                                            : exposure_host(a.exposure_function, a.exposure_optimum, a.exposure_width, (double)g / maxv);
        lut[(size_t)g] = (float)(p->w_exposure * wt);
    }
}

}  // namespace host
}  // namespace enbl

namespace {

struct WeightScratch {
    float* lut = nullptr;
    void* gray = nullptr;
    void* grad = nullptr;
    double* tab = nullptr;
    void* ent = nullptr;
    long pitch = 0;   // elements, shared by gray/grad/ent
    size_t plane = 0; // elements between the images' gray/grad planes
    // contrast by edges: kernels (smoothing, second derivative, LCE smoothing
    // at ks, kd, kl), two real scratch planes, the sharpened gray plane and
    // the K local standard deviations of the fill-in functor
    const double *ks = nullptr, *kd = nullptr, *kl = nullptr;
    int rs = 0, rd = 0, rl = 0;
    void *ta = nullptr, *tb = nullptr, *lce = nullptr, *lc = nullptr;
};

// Host tables and device scratch for K images of one size (ExposureFunctor
// table, K gray and K gradient planes, entropy tables).
int prepare_weights(enbl_context* ctx, int itype, int w, int h, int K, const enbl_enfuse_params* p,
                    const WeightArgs& a, WeightScratch* ws)
{
    const long pitch = (long)round_up((size_t)w, 64);
    ws->pitch = pitch;
    const size_t npx = (size_t)pitch * h;
    ws->plane = npx;
    int rc;
    if (p->w_exposure > 0.0 && itype != ENBL_F32) {
        std::vector<float> lut;
        exposure_lut(itype, p, a, lut);
        const int n = (int)lut.size();
        void* d;
        rc = scratch(ctx, 0, (size_t)n * sizeof(float), &d);
        if (rc) return rc;
        rc = hip_check(hipMemcpyAsync(d, lut.data(), (size_t)n * sizeof(float), hipMemcpyHostToDevice, ctx->stream), "exposure table");
        if (rc) return rc;
        rc = hip_check(hipStreamSynchronize(ctx->stream), "exposure table");
        if (rc) return rc;
        ws->lut = (float*)d;
    }
    if (p->w_contrast > 0.0) {
        rc = scratch(ctx, 1, npx * 4 * (size_t)K, &ws->gray);
        if (!rc) rc = scratch(ctx, 2, npx * 4 * (size_t)K, &ws->grad);
        if (rc) return rc;
    }
    if (p->w_contrast > 0.0 && p->edge_scale > 0.0) {
        const std::vector<double> ks = vigra_kernel(p->edge_scale, 0), kd = vigra_kernel(p->edge_scale, 2);
        std::vector<double> kl;
        if (p->lce_scale > 0.0) kl = vigra_kernel(p->lce_scale, 0);
        std::vector<double> all(ks);
        all.insert(all.end(), kd.begin(), kd.end());
        all.insert(all.end(), kl.begin(), kl.end());
        void* d;
        rc = scratch(ctx, 27, all.size() * sizeof(double), &d);
        if (rc) return rc;
        rc = hip_check(hipMemcpyAsync(d, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream),
                       "edge kernels");
        if (!rc) rc = hip_check(hipStreamSynchronize(ctx->stream), "edge kernels");
        if (rc) return rc;
        ws->ks = (const double*)d;
        ws->kd = ws->ks + ks.size();
        ws->kl = p->lce_scale > 0.0 ? ws->kd + kd.size() : nullptr;
        ws->rs = (int)(ks.size() / 2);
        ws->rd = (int)(kd.size() / 2);
        ws->rl = (int)(kl.size() / 2);
        const size_t es_r = itype == ENBL_F32 ? 4 : 8;
        void* t;
        rc = scratch(ctx, 28, 2 * npx * es_r, &t);
        if (rc) return rc;
        ws->ta = t;
        ws->tb = (char*)t + npx * es_r;
        rc = scratch(ctx, 29, npx * 4 * (size_t)(1 + K), &t);
        if (rc) return rc;
        ws->lce = t;
        ws->lc = (char*)t + npx * 4;
    }
    if (p->w_entropy > 0.0) {
        const int full = p->entropy_window * p->entropy_window;
        // pent, plog, the general table, then the full-window sum of `full`
        // one-sample terms in order (a window of distinct keys)
        std::vector<double> tab((size_t)(2 * (full + 1) + (full + 1) * (full + 1) + 1), 0.0);
        double* pent = tab.data();
        double* plog = pent + (full + 1);
        double* pgen = plog + (full + 1);
        for (int i = 1; i <= full; ++i) {
            const double pr = (double)i / (double)full;
            plog[i] = log((double)i);
            pent[i] = pr * log(pr);
        }
        for (int t = 1; t <= full; ++t)
            for (int k = 1; k <= t; ++k) {
                const double pr = k / (double)t;
                pgen[t * (full + 1) + k] = pr * log(pr);
            }
        double sd = 0.0;
        for (int i = 0; i < full; ++i) sd += pent[1];
        tab[(size_t)(2 * (full + 1) + (full + 1) * (full + 1))] = sd;
        void* d;
        rc = scratch(ctx, 3, tab.size() * sizeof(double), &d);
        if (rc) return rc;
        rc = hip_check(hipMemcpyAsync(d, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream), "entropy table");
        if (!rc) rc = hip_check(hipStreamSynchronize(ctx->stream), "entropy table");
        if (rc) return rc;
        ws->tab = (double*)d;
        rc = scratch(ctx, 4, npx * 4, &ws->ent);
        if (rc) return rc;
    }
    return ENBL_OK;
}

// Contrast by Laplacian-of-Gaussian edges (src/enfuse.h:1020-1093) of all K
// gray planes: the local standard deviations first when the fill-in functor
// needs them, then per image the optional sharpening (gaussianSharpening)
// and the LoG with its clamping or fill-in functor into gradient plane m.
int edge_contrast_all(enbl_context* ctx, int itype, int w, int h, int K, const uint8_t* const* alphas, const long* aps,
                      const enbl_enfuse_params* p, const WeightScratch& ws)
{
    const hipStream_t s = ctx->stream;
    const size_t es = 4;
    const double es_r = itype == ENBL_F32 ? 4.0 : 8.0;
    const double mc = p->min_curvature;
    int rc;
    if (mc > 0.0) {
        if (itype == ENBL_F32) {
            rc = launch_fill(ws.lc, ws.plane * es * (size_t)K, s);
            if (!rc) rc = launch_status("fill");
            if (rc) return rc;
        }
        Timed t(ctx, "stddev", (double)K * w * h * 9.0);
        rc = launch_stddev(itype, ws.gray, ws.pitch, ws.plane, alphas, aps, K, w, h, p->contrast_window, ws.lc,
                           ws.pitch, ws.plane, s);
        if (rc) return fail(rc, "localStdDevIf: unsupported window");
        rc = launch_status("stddev");
        if (rc) return rc;
    }
    const double smax = itype == ENBL_U8 ? 255.0 : itype == ENBL_U16 ? 65535.0 : (double)FLT_MAX;
    for (int m = 0; m < K; ++m) {
        const void* gray = (const char*)ws.gray + (size_t)m * ws.plane * es;
        EdgePlanes e;
        e.src = gray; e.spitch = ws.pitch; e.tA = ws.ta; e.tB = ws.tb; e.tpitch = ws.pitch; e.w = w; e.h = h;
        if (ws.kl) {
            {
                Timed t(ctx, "edge_rows", (double)w * h * (4.0 + es_r));
                rc = launch_edge_rows(itype, e, ws.kl, ws.rl, nullptr, 0, s);
                if (!rc) rc = launch_status("edge_rows");
                if (rc) return rc;
            }
            Timed t(ctx, "edge_sharpen", (double)w * h * (es_r + 8.0));
            rc = launch_edge_sharpen(itype, e, ws.kl, ws.rl, p->lce_factor, gray, ws.pitch, ws.lce, ws.pitch, s);
            if (!rc) rc = launch_status("edge_sharpen");
            if (rc) return rc;
            e.src = ws.lce;
        }
        {
            Timed t(ctx, "edge_rows", (double)w * h * (4.0 + 2.0 * es_r));
            rc = launch_edge_rows(itype, e, ws.kd, ws.rd, ws.ks, ws.rs, s);
            if (!rc) rc = launch_status("edge_rows");
            if (rc) return rc;
        }
        Timed t(ctx, "edge_log", (double)w * h * (2.0 * es_r + 1.0 + 4.0 + (mc > 0.0 ? 4.0 : 0.0)));
        rc = launch_edge_log(itype, e, ws.ks, ws.rs, ws.kd, ws.rd, alphas[m], aps[m],
                             (const char*)ws.lc + (size_t)m * ws.plane * es, ws.pitch, mc, smax,
                             (char*)ws.grad + (size_t)m * ws.plane * es, ws.pitch, s);
        if (!rc) rc = launch_status("edge_log");
        if (rc) return rc;
    }
    return 0;
}

// Gray images and local standard deviations (src/enfuse.h:1099-1120) of
// all K images: gradient plane m of ws.grad belongs to image m.
// Gray planes of all images (and, for float windows the tile kernel does not
// take, the cleared gradient planes).
int contrast_gray(enbl_context* ctx, int itype, int w, int h, int K, const void* const* imgs, const long* ips,
                  const enbl_enfuse_params* p, const WeightScratch& ws)
{
    const hipStream_t s = ctx->stream;
    const size_t es = 4;   // int32 gray of integer images, float gray of float images
    int rc;
    for (int m = 0; m < K; ++m) {
        Timed t(ctx, "gray", (double)w * h * (3.0 * ptype_size(itype) + 4.0));
        rc = launch_gray(itype, imgs[m], ips[m], w, h, (char*)ws.gray + m * ws.plane * es, ws.pitch, gray_proj(p), s);
        if (!rc) rc = launch_status("gray");
        if (rc) return rc;
    }
    if (p->edge_scale <= 0.0 && itype == ENBL_F32 && p->contrast_window != 3 && p->contrast_window != 5 &&
        p->contrast_window != 7) {
        // the row kernel of other windows writes the interior under alpha
        // only (the tile kernel of 3, 5 and 7 writes every pixel)
        rc = launch_fill(ws.grad, ws.plane * es * (size_t)K, s);
        if (!rc) rc = launch_status("fill");
        if (rc) return rc;
    }
    return ENBL_OK;
}

// Local standard deviations of images m0 .. m1-1 from their gray planes.
int contrast_stddev(enbl_context* ctx, int itype, int w, int h, int m0, int m1, const uint8_t* const* alphas,
                    const long* aps, const enbl_enfuse_params* p, const WeightScratch& ws)
{
    if (m1 <= m0) return ENBL_OK;
    const size_t es = 4;
    const size_t off = (size_t)m0 * ws.plane * es;
    Timed t(ctx, "stddev", (double)(m1 - m0) * w * h * 9.0);
    const int rc = launch_stddev(itype, (const char*)ws.gray + off, ws.pitch, ws.plane, alphas + m0, aps + m0, m1 - m0, w,
                                 h, p->contrast_window, (char*)ws.grad + off, ws.pitch, ws.plane, ctx->stream);
    if (rc) return fail(rc, "localStdDevIf: unsupported window");
    return launch_status("stddev");
}

int contrast_all(enbl_context* ctx, int itype, int w, int h, int K, const void* const* imgs, const long* ips,
                 const uint8_t* const* alphas, const long* aps, const enbl_enfuse_params* p, const WeightScratch& ws)
{
    int rc = contrast_gray(ctx, itype, w, h, K, imgs, ips, p, ws);
    if (rc) return rc;
    if (p->edge_scale > 0.0) return edge_contrast_all(ctx, itype, w, h, K, alphas, aps, p, ws);
    return contrast_stddev(ctx, itype, w, h, 0, K, alphas, aps, p, ws);
}

// Weights of image m into mask (its gradient plane made by contrast_all);
// norm += mask; union alpha updated.
int image_weights(enbl_context* ctx, int itype, int w, int h, int m, const void* img, long ip, const uint8_t* alpha,
                  long ap, const enbl_enfuse_params* p, const WeightArgs& a, const WeightScratch& ws,
                  float* mask, long mp, float* norm, long np_, uint8_t* ua, long up)
{
    const hipStream_t s = ctx->stream;
    int rc;
    const void* grad = ws.grad ? (const void*)((const char*)ws.grad + (size_t)m * ws.plane * 4) : nullptr;
    const int n = p->entropy_window;
    if (p->w_entropy > 0.0 && !p->entropy_cutoff && (n == 3 || n == 5 || n == 7)) {
        // the entropy kernel holds the pixels and makes the whole weight
        Timed t(ctx, "entropy_weights", (double)w * h * (3.0 * ptype_size(itype) + 1.0 + 4.0 + 4.0 + 8.0 + 1.0));
        rc = launch_entropy_weights(itype, img, ip, alpha, ap, w, h, n, ws.tab, a, ws.lut, grad, ws.pitch, mask, mp,
                                    norm, np_, ua, up, s);
        if (rc) return fail(rc, "localEntropyIf: unsupported window");
        return launch_status("entropy_weights");
    }
    if (p->w_entropy > 0.0) {
        const void* src = img;
        long sp = ip;
        if (p->entropy_cutoff) {
            void* tr;
            rc = scratch(ctx, 30, (size_t)w * h * 3 * ptype_size(itype), &tr);
            if (rc) return rc;
            Timed t(ctx, "entropy_clamp", (double)w * h * 6.0 * ptype_size(itype));
            rc = launch_entropy_clamp(itype, img, ip, w, h, p->entropy_lower_cutoff, p->entropy_upper_cutoff, tr, s);
            if (!rc) rc = launch_status("entropy_clamp");
            if (rc) return rc;
            src = tr;
            sp = 3L * w;
        }
        Timed t(ctx, "entropy", (double)w * h * (3.0 * ptype_size(itype) + 1.0 + ptype_size(itype)));
        rc = launch_entropy(itype, src, sp, alpha, ap, w, h, p->entropy_window, ws.tab, ws.ent, ws.pitch, s);
        if (rc) return fail(rc, "localEntropyIf: unsupported window");
        rc = launch_status("entropy");
        if (rc) return rc;
    }
    Timed t(ctx, "weights", (double)w * h * (3.0 * ptype_size(itype) + 1.0 + 4.0 + 8.0 + 1.0));
    rc = launch_weights(itype, img, ip, alpha, ap, w, h, a, ws.lut, grad, ws.pitch, ws.ent, ws.pitch,
                        mask, mp, norm, np_, ua, up, s);
    if (rc) return fail(rc, "enfuse weights: unsupported image type");
    return launch_status("weights");
}

}  // namespace

namespace enbl {
namespace host {

// The first pass of enbl_enfuse for one stack (src/enfuse.h:1277-1388 and
// the normalisation of :1529-1546): every image's weight, the norm and the
// union alpha, then the fixed-point mask level 0 of every image.
int enfuse_stack_masks(enbl_context* ctx, int itype, int w, int h, int K, const void* const* imgs, const long* ips,
                       const uint8_t* const* alphas, const long* aps, const enbl_enfuse_params* p, uint8_t* ua,
                       long uap, void* const* mask0, long mpitch)
{
    const WeightArgs a = weight_args(p);
    WeightScratch ws;
    int rc = prepare_weights(ctx, itype, w, h, K, p, a, &ws);
    if (rc) return rc;
    const hipStream_t s = ctx->stream;
    const long pitch = ws.pitch;
    const size_t plane = (size_t)pitch * h;
    void *norm, *masks;
    rc = scratch(ctx, 5, plane * 4, &norm);
    if (!rc) rc = scratch(ctx, 7, plane * 4 * (size_t)K, &masks);
    if (!rc) rc = launch_fill(norm, plane * 4, s);
    if (!rc) rc = launch_fill(ua, (size_t)uap * h, s);
    if (!rc && p->w_contrast > 0.0) rc = contrast_all(ctx, itype, w, h, K, imgs, ips, alphas, aps, p, ws);
    for (int m = 0; m < K && !rc; ++m)
        rc = image_weights(ctx, itype, w, h, m, imgs[m], ips[m], alphas[m], aps[m], p, a, ws,
                           (float*)masks + plane * m, pitch, (float*)norm, pitch, ua, uap);
    if (!rc && p->hard_mask) {
        Timed t(ctx, "hard_mask", (double)w * h * 8.0 * K);
        rc = launch_hard_mask((float*)masks, pitch, plane, w, h, K, s);
        if (!rc) rc = launch_status("hard_mask");
    }
    const int mt = mask_ptype_for(itype);
    const int mfrac = itype == ENBL_U8 ? 7 : 15;
    for (int m = 0; m < K && !rc; ++m) {
        Timed t(ctx, "normalize", (double)w * h * (8.0 + ptype_size(mt)));
        rc = launch_normalize(mt, (const float*)masks + plane * m, pitch, (const float*)norm, pitch, w, h,
                              (float)(255 / K), (double)(1u << mfrac), p->hard_mask, mask0[m], mpitch, s);
        if (!rc) rc = launch_status("normalize");
    }
    return rc;
}

}  // namespace host
}  // namespace enbl

extern "C" {

int enbl_enfuse_mask(enbl_context* ctx, int itype, int w, int h, const void* img, long ip,
                     const uint8_t* alpha, long ap, const enbl_enfuse_params* p, float* mask, long mp)
{
    if (!ctx || !img || !alpha || !mask) return fail(ENBL_E_ARG, "enbl_enfuse_mask: null argument");
    int rc = check_params(itype, w, h, p);
    if (rc) return rc;
    const WeightArgs a = weight_args(p);
    WeightScratch ws;
    rc = prepare_weights(ctx, itype, w, h, 1, p, a, &ws);
    if (!rc && p->w_contrast > 0.0) rc = contrast_all(ctx, itype, w, h, 1, &img, &ip, &alpha, &ap, p, ws);
    if (rc) return rc;
    // norm and union alpha are by-products here; park them in scratch.
    void *norm, *ua;
    const long pitch = ws.pitch;
    rc = scratch(ctx, 5, (size_t)pitch * h * 4, &norm);
    if (!rc) rc = scratch(ctx, 6, (size_t)pitch * h, &ua);
    if (rc) return rc;
    rc = launch_fill(norm, (size_t)pitch * h * 4, ctx->stream);
    if (rc) return rc;
    return image_weights(ctx, itype, w, h, 0, img, ip, alpha, ap, p, a, ws, mask, mp, (float*)norm, pitch,
                         (uint8_t*)ua, pitch);
}

int enbl_enfuse(enbl_context* ctx, int itype, int w, int h, int K, const void* const* imgs, const long* ips,
                const uint8_t* const* alphas, const long* aps, const enbl_enfuse_params* p, int levels, int wrap,
                void* out, long op, uint8_t* oa, long oap, int dither, uint64_t seed)
{
    if (!ctx || !imgs || !ips || !alphas || !aps || !out || !oa) return fail(ENBL_E_ARG, "enbl_enfuse: null argument");
    if (K < 1) return fail(ENBL_E_ARG, "enbl_enfuse: no images");
    for (int m = 0; m < K; ++m)
        if (!imgs[m] || !alphas[m]) return fail(ENBL_E_ARG, "enbl_enfuse: null image %d", m);
    int rc = check_params(itype, w, h, p);
    if (rc) return rc;
    if (levels < 1 || levels > 29) return fail(ENBL_E_PRECONDITION, "filterHalfWidth: levels outside of permissible range");
    const int cs = ctx->colorspace;
    rc = check_blend_space(cs, itype, 3);
    if (rc) return rc;
    const WeightArgs a = weight_args(p);
    WeightScratch ws;
    rc = prepare_weights(ctx, itype, w, h, K, p, a, &ws);
    if (rc) return rc;
    const hipStream_t s = ctx->stream;
    const long pitch = ws.pitch;
    const size_t plane = (size_t)pitch * h;
    // In a CIELAB / CIELUV blend space the weights come from the RGB images
    // and the pyramids from the images converted into scaled blend-space
    // images (ConvertVectorTo{Lab,Luv}PyramidFunctor); the result goes back
    // to RGB under the union alpha.
    std::vector<const void*> pimgs(imgs, imgs + K);
    std::vector<long> pips(ips, ips + K);
    void* csout = nullptr;
    // CIECAM02: the JCh level 0 (with fraction bits) goes straight into each
    // image's pyramid and the result's level 0 back to RGB.
    const bool jch = cs == ENBL_COLORSPACE_CIECAM;
    if (cs == ENBL_COLORSPACE_CIELAB || cs == ENBL_COLORSPACE_CIELUV) {
        const size_t cplane = (size_t)w * h * 3 * (itype == ENBL_U8 ? 1 : 2);
        void* cbuf;
        rc = scratch(ctx, 33, cplane * (size_t)K, &cbuf);
        if (!rc) rc = scratch(ctx, 34, cplane, &csout);
        if (rc) return rc;
        for (int m = 0; m < K; ++m) {
            pimgs[(size_t)m] = (char*)cbuf + cplane * m;
            pips[(size_t)m] = 3L * w;
            rc = rgb_to_blend_space(ctx, cs, itype, w, h, imgs[m], ips[m], (void*)pimgs[(size_t)m], 3L * w);
            if (rc) return rc;
        }
    }
    void *norm, *ua, *masks;
    rc = scratch(ctx, 5, plane * 4, &norm);
    if (!rc) rc = scratch(ctx, 6, plane, &ua);
    if (!rc) rc = scratch(ctx, 7, plane * 4 * (size_t)K, &masks);
    if (rc) return rc;
    rc = launch_fill(norm, plane * 4, s);
    if (!rc) rc = launch_fill(ua, plane, s);
    if (rc) return rc;

    const int pt = image_ptype_for(itype);
    const int mt = mask_ptype_for(itype);
    const int ifrac = itype == ENBL_F32 ? 0 : 7;
    const int mfrac = itype == ENBL_U8 ? 7 : 15;
    // Every image's Laplacian pyramid and normalised mask pyramid are kept
    // (context cache), then each level merges up to kMergeMax images per
    // pass in image order, so the result level is read and written once per
    // group instead of once per image.
    enbl_pyramid* res = nullptr;
    std::vector<enbl_pyramid*> lps((size_t)K, nullptr), mgps((size_t)K, nullptr);
    rc = cached_pyramid(ctx, kRoleResult, pt, 3, w, h, levels, 0, &res);
    if (rc) return rc;
    for (int m = 0; m < K; ++m) {
        rc = cached_pyramid(ctx, kRoleImage, pt, 3, w, h, levels, m, &lps[(size_t)m]);
        if (!rc) rc = cached_pyramid(ctx, kRoleMask, mt, 1, w, h, levels, m, &mgps[(size_t)m]);
        if (rc) return rc;
    }
    // The images' Laplacian pyramids need the images alone: without
    // wraparound they are built on the auxiliary stream while the weights
    // are computed on the main one (memory-bound passes beside the
    // arithmetic-bound weight kernels); the main stream waits for them
    // before the merge, and on every error path.
    auto jch_laplacian = [&](int m) {
        enbl_pyramid* lp = lps[(size_t)m];
        const Level& L0 = lp->lv[0];
        int r = rgb_to_jch(ctx, itype, w, h, imgs[m], ips[m], level_ptr(lp, 0), L0.pitch, L0.plane);
        if (!r) r = gaussian_levels(ctx, lp, wrap, alphas[m], aps[m]);
        if (!r) r = laplacian_levels(ctx, lp, wrap);
        return r;
    };
    struct AuxJoin {
        enbl_context* ctx;
        bool active = false;
        ~AuxJoin() { if (active) (void)join_aux(ctx); }
    } aux{ctx};
    // Level 0 on the fly (float images): when level 1 is a streaming level,
    // each image's level 0 is only reduced into its Gaussian level 1 (kept
    // apart) while its log transform is written once, and the merge of level
    // 0, the last collapse step and the conversion into the output are one
    // pass that forms every Laplacian level 0 from those two: no Laplacian
    // level 0 is stored.  8- and 16-bit stacks keep the stored integer level
    // 0, which costs less than forming it on the fly (64.3 against 65.1 ms
    // for the 64-stack batch run stack by stack).
    const bool fuse0 = itype == ENBL_F32 && !wrap && cs == ENBL_COLORSPACE_IDENTITY && levels >= 2 && K <= kL0Max &&
                       enfuse_level0_count(lps[0], alphas[0]) >= 1;
    std::vector<void*> g1s;
    if (fuse0) {
        const Level& D = lps[0]->lv[1];
        const size_t gbytes = round_up((size_t)D.plane * 3 * ptype_size(pt), 256);
        void* g;
        rc = scratch(ctx, 37, gbytes * (size_t)K, &g);
        if (rc) return rc;
        for (int m = 0; m < K; ++m) g1s.push_back((char*)g + gbytes * (size_t)m);
    }
    if (!wrap && jch) {
        // (level by level on the main stream: these reduces share the
        // context's alpha scratch with the mask pyramids)
        for (int m = 0; m < K && !rc; ++m) rc = jch_laplacian(m);
        if (rc) return rc;
    }
    // Pass 1: weights, norm, union alpha (src/enfuse.h:1277-1388); the
    // contrast of all images first, so that their local standard deviations
    // share launches.  Beside the image pyramids (auxiliary stream), the
    // standard deviations of the second half of the images run on the
    // auxiliary stream too, ahead of the pyramids, while the first half's
    // weights are made; the second half's weights wait for them.
    const bool sd_split = !wrap && !jch && p->w_contrast > 0.0 && p->edge_scale <= 0.0 && K >= 2;
    const int half = sd_split ? (K + 1) / 2 : K;
    bool forked = false;
    if (sd_split) {
        rc = contrast_gray(ctx, itype, w, h, K, imgs, ips, p, ws);
        if (rc) return rc;
    }
    if (!wrap && !jch) {
        rc = fork_aux(ctx);
        if (rc) return rc;
        aux.active = true;
        forked = ctx->forked_from != nullptr;
        if (sd_split) {
            rc = contrast_stddev(ctx, itype, w, h, half, K, alphas, aps, p, ws);
            if (!rc && forked) rc = hip_check(hipEventRecord(ctx->aux_mid, ctx->stream), "hipEventRecord");
        }
        if (!rc)
            rc = enfuse_image_pyramids(ctx, itype, ifrac, K, pimgs.data(), pips.data(), alphas, aps, lps.data(),
                                       fuse0 ? g1s.data() : nullptr);
        const int rc2 = end_aux(ctx);
        if (!rc) rc = rc2;
        if (rc) return rc;
    }
    if (sd_split) {
        rc = contrast_stddev(ctx, itype, w, h, 0, half, alphas, aps, p, ws);
        if (rc) return rc;
    } else if (p->w_contrast > 0.0) {
        rc = contrast_all(ctx, itype, w, h, K, imgs, ips, alphas, aps, p, ws);
        if (rc) return rc;
    }
    for (int m = 0; m < K; ++m) {
        if (sd_split && forked && m == half) {
            rc = hip_check(hipStreamWaitEvent(ctx->stream, ctx->aux_mid, 0), "hipStreamWaitEvent");
            if (rc) return rc;
        }
        rc = image_weights(ctx, itype, w, h, m, imgs[m], ips[m], alphas[m], aps[m], p, a, ws,
                           (float*)masks + plane * m, pitch, (float*)norm, pitch, (uint8_t*)ua, pitch);
        if (rc) return rc;
    }
    if (p->hard_mask) {
        Timed t(ctx, "hard_mask", (double)w * h * 8.0 * K);
        rc = launch_hard_mask((float*)masks, pitch, plane, w, h, K, s);
        if (!rc) rc = launch_status("hard_mask");
        if (rc) return rc;
    }

    // Pass 2: per input, Laplacian pyramid, normalised mask Gaussian pyramid
    // over the union alpha, multiply and accumulate (src/enfuse.h:1497-1590).
    const float fallback = (float)(255 / K);
    const double scale = (double)(1u << mfrac);
    const double maxpyr = (double)(255 << mfrac);
    auto normalize = [&](int m) {
        enbl_pyramid* mgp = mgps[(size_t)m];
        const Level& L0 = mgp->lv[0];
        Timed t(ctx, "normalize", (double)w * h * (8.0 + ptype_size(mt)));
        int r = launch_normalize(mt, (const float*)masks + plane * m, pitch, (const float*)norm, pitch, w, h,
                                 fallback, scale, p->hard_mask, level_ptr(mgp, 0), L0.pitch, s);
        return r ? r : launch_status("normalize");
    };
    std::vector<const float*> wts;
    for (int m = 0; m < K; ++m) wts.push_back((const float*)masks + plane * m);
    const MaskLevel0 ml0{wts.data(), pitch, (const float*)norm, pitch, scale, fallback, p->hard_mask};
    if (!wrap) {
        // mask level 0 of every image (formed on the fly from the weights and
        // the norm where level 0 is merged on the fly), the mask pyramids with
        // their small levels in shared launches, then the image pyramids' end
        if (!fuse0)
            for (int m = 0; m < K && !rc; ++m) rc = normalize(m);
        if (!rc) rc = enfuse_mask_pyramids(ctx, K, mgps.data(), (const uint8_t*)ua, pitch, fuse0 ? &ml0 : nullptr);
        if (rc) return rc;
        aux.active = false;
        rc = join_aux(ctx);
        if (rc) return rc;
    } else {
        for (int m = 0; m < K; ++m) {
            rc = jch ? jch_laplacian(m)
                     : enbl_laplacian_pyramid(ctx, wrap, itype, ifrac, pimgs[(size_t)m], pips[(size_t)m], alphas[m],
                                              aps[m], lps[(size_t)m]);
            if (!rc) rc = normalize(m);
            if (!rc) rc = gaussian_levels(ctx, mgps[(size_t)m], wrap, (const uint8_t*)ua, pitch);
            if (rc) return rc;
        }
    }
    for (int l = fuse0 ? 1 : 0; l < levels; ++l) {
        const Level& L = lps[0]->lv[l];
        const Level& G = mgps[0]->lv[l];
        const Level& R = res->lv[l];
        for (int m0 = 0; m0 < K; m0 += kMergeMax) {
            MergeMany a;
            memset(&a, 0, sizeof a);
            a.n = std::min(kMergeMax, K - m0);
            for (int k = 0; k < a.n; ++k) {
                a.lp[k] = level_ptr(lps[(size_t)(m0 + k)], l);
                a.g[k] = level_ptr(mgps[(size_t)(m0 + k)], l);
            }
            a.lpp = L.pitch; a.lpl = L.plane; a.gpp = G.pitch;
            a.res = level_ptr(res, l); a.rp = R.pitch; a.rl = R.plane;
            a.maxpyr = maxpyr; a.w = L.w; a.h = L.h; a.first = m0 == 0;
            // credited: the image-by-image merges it replaces; own: each
            // image's level and mask read once, the result written once (and
            // read once for later groups)
            const double esp = 3.0 * ptype_size(pt), px = (double)L.w * L.h;
            const double credited = px * (a.n * (esp + ptype_size(mt)) + esp * (2.0 * a.n - (m0 == 0 ? 1.0 : 0.0)));
            const double own = px * (a.n * (esp + ptype_size(mt)) + esp * (m0 == 0 ? 1.0 : 2.0));
            Timed t(ctx, "merge", credited, own);
            rc = launch_merge_many(pt, mt, a, s);
            if (!rc) rc = launch_status("merge");
            if (rc) return rc;
        }
    }
    rc = collapse_levels(ctx, wrap, res, fuse0 ? 1 : 0);
    if (rc) return rc;
    if (fuse0) {
        L0FuseArgs a;
        memset(&a, 0, sizeof a);
        const Level& C1 = lps[0]->lv[1];
        const Level& R1 = res->lv[1];
        a.K = K; a.fw = w; a.fh = h; a.cw = C1.w; a.ch = C1.h;
        const bool g0 = itype == ENBL_F32;   // the converted level 0 is in each image pyramid's level 0
        for (int m = 0; m < K; ++m) {
            a.img[m] = imgs[m]; a.ipitch[m] = ips[m];
            a.g0[m] = g0 ? level_ptr(lps[(size_t)m], 0) : nullptr;
            a.g1[m] = g1s[(size_t)m];
            a.wt[m] = wts[(size_t)m];
        }
        a.wpitch = pitch; a.norm = (const float*)norm; a.npitch = pitch;
        a.mscale = scale; a.fallback = fallback; a.hard = p->hard_mask;
        a.g0pitch = lps[0]->lv[0].pitch; a.g0plane = lps[0]->lv[0].plane;
        a.gpitch = C1.pitch; a.gplane = C1.plane;
        a.rc = level_ptr(res, 1); a.rcpitch = R1.pitch; a.rcplane = R1.plane;
        a.ualpha = (const uint8_t*)ua; a.upitch = pitch;
        a.out = out; a.opitch = op;
        a.oalpha = oa; a.oapitch = oap;
        a.maxpyr = maxpyr; a.frac = ifrac; a.dither = dither; a.seed = seed;
        // own bytes: per image its pixels, mask level 0 and Gaussian level 1
        // read once; result level 1 read once; union alpha read, output image
        // and alpha written
        const double esP = 3.0 * ptype_size(pt), es_i = 3.0 * ptype_size(itype);
        const double n0 = (double)w * h, n1 = (double)C1.w * C1.h;
        char name[64];
        snprintf(name, sizeof name, "merge_level0_%s_c3", pname(pt));
        Timed t(ctx, name, K * (n0 * ((g0 ? esP : es_i) + 4.0) + n1 * esP) + n1 * esP + n0 * (2.0 + 4.0 + es_i));
        rc = launch_enfuse_level0(itype, a, s);
        if (rc) return fail(rc, "level-0 merge: unsupported type");
        return launch_status(name);
    }
    if (jch) {
        const Level& R0 = res->lv[0];
        rc = jch_to_rgb_if(ctx, itype, w, h, level_ptr(res, 0), R0.pitch, R0.plane, (const uint8_t*)ua, pitch, out, op);
    } else if (cs != ENBL_COLORSPACE_IDENTITY) {
        rc = enbl_copy_from_pyramid_if(ctx, res, 0, ifrac, (const uint8_t*)ua, pitch, csout, itype, 3L * w, dither,
                                       seed);
        if (!rc) rc = blend_space_to_rgb_if(ctx, cs, itype, w, h, csout, 3L * w, (const uint8_t*)ua, pitch, out, op);
    } else {
        rc = enbl_copy_from_pyramid_if(ctx, res, 0, ifrac, (const uint8_t*)ua, pitch, out, itype, op, dither, seed);
    }
    if (rc) return rc;
    return hip_check(hipMemcpy2DAsync(oa, (size_t)oap, ua, (size_t)pitch, (size_t)w, (size_t)h, hipMemcpyDeviceToDevice, s),
                     "union alpha");
}

}  // extern "C"
