// CPU restatement of enfuse's weighting and fusion (TEST INFRASTRUCTURE
// ONLY: checker of the HIP path and bench.py's CPU baseline).
//
// Follows src/enfuse.h:64-1212 (weights) and :1217-1609 (enfuseMain) of the
// reference, with the default "average" grayscale projector (src/mga.h:188-193),
// soft masks and the identity colour space.  Parallel structure as in the
// reference: pointwise stages and the local standard deviation run over rows
// with OpenMP, the local entropy is one sequential column sweep.
#include <float.h>

#include <map>

#include "oracle_core.h"

using namespace orc;

namespace {

// ---- grayscale projectors, src/mga.h:60-250 ---------------------------------
// Integer channels promote to double; float channels promote to float
// (RealPromote<float> = float), so their sum is a float sum.  Every projector
// result lies inside the channel range, so rounding it to the scalar type and
// then to a wider result type R equals rounding it to R directly.
//   average    (r + g + b) / 3
//   l-star     max * L* / 100 of linear RGB (VIGRA RGB2LabFunctor: Y of the
//              sRGB/D65 matrix, L* = Y < 216/24389 ? 24389/27 Y : 116 Y^(1/3) - 16)
//   pl-star    the same after the gamma 1/0.45 (VIGRA RGBPrime2LabFunctor)
//   lightness  (min + max) / 2      value  max      anti-value  min
//   luminance  RGBValue::luminance(): 0.3 r + 0.59 g + 0.11 b, cast to the channel type
//   channel-mixer  wr r + wg g + wb b with the weights divided by their sum
// VIGRA's colour functors are not vendored; their constants are restated.
template <class T> struct RealOf { typedef double R; };
template <> struct RealOf<float> { typedef float R; };

template <class T> double scalar_max();
template <> double scalar_max<uint8_t>() { return 255.0; }
template <> double scalar_max<uint16_t>() { return 65535.0; }
template <> double scalar_max<float>() { return (double)FLT_MAX; }

template <class T> T to_channel(double v) { return fromRealPromote<T>(v); }
template <> float to_channel<float>(double v) { return (float)v; }

double lstar_y(double r, double g, double b)
{
    const double Y = 0.212671 * r + 0.715160 * g + 0.072169 * b;
    return Y < 216.0 / 24389.0 ? 24389.0 / 27.0 * Y : 116.0 * pow(Y, 1.0 / 3.0) - 16.0;
}

double gamma_correct(double v, double gamma) { return v < 0.0 ? -pow(-v, gamma) : pow(v, gamma); }

template <class T>
T project(const T* p, const orc_enfuse_params* prm)
{
    typedef typename RealOf<T>::R RP;
    const T r = p[0], g = p[1], b = p[2];
    switch (prm->gray_projector) {
    case ORC_GRAY_LSTAR:
    case ORC_GRAY_PLSTAR: {
        const double mx = scalar_max<T>();
        double rr = (double)r / mx, gg = (double)g / mx, bb = (double)b / mx;
        if (prm->gray_projector == ORC_GRAY_PLSTAR) {
            rr = gamma_correct(rr, 1.0 / 0.45);
            gg = gamma_correct(gg, 1.0 / 0.45);
            bb = gamma_correct(bb, 1.0 / 0.45);
        }
        return to_channel<T>(mx * (lstar_y(rr, gg, bb) / 100.0));
    }
    case ORC_GRAY_LIGHTNESS: {
        const T mn = std::min(r, std::min(g, b)), mx = std::max(r, std::max(g, b));
        return to_channel<T>((double)(mn + mx) / 2.0);
    }
    case ORC_GRAY_VALUE: return std::max(r, std::max(g, b));
    case ORC_GRAY_ANTI_VALUE: return std::min(r, std::min(g, b));
    case ORC_GRAY_LUMINANCE: return to_channel<T>(0.3 * (double)r + 0.59 * (double)g + 0.11 * (double)b);
    case ORC_GRAY_MIXER: {
        const double sum = prm->mixer_red + prm->mixer_green + prm->mixer_blue;
        const double wr = prm->mixer_red / sum, wg = prm->mixer_green / sum, wb = prm->mixer_blue / sum;
        return to_channel<T>(wr * (double)(RP)r + wg * (double)(RP)g + wb * (double)(RP)b);
    }
    default:
        return to_channel<T>((double)((RP)r + (RP)g + (RP)b) / 3.0);
    }
}

template <class T, class R> struct Gray {
    static R get(const T* p, const orc_enfuse_params* prm) { return (R)project<T>(p, prm); }
};

template <class T> struct Max;
template <> struct Max<uint8_t> { static double v() { return 255.0; } };
template <> struct Max<uint16_t> { static double v() { return 65535.0; } };
template <> struct Max<float> { static double v() { return (double)FLT_MAX; } };

// ---- exposure weight functions, src/exposure_weight.h:44-120 -------------
const double kFwhmGaussian = 2.3548200450309493820231386529193992755;

struct Exposure {
    int fn;
    double opt, width;
    Exposure(int f, double o, double w) : fn(f), opt(o), width(w) {
        static const double fwhm[] = {kFwhmGaussian, 2.8284271247461900976033774484193961571,
                                      2.0943951023931954923084289221863352561, M_PI,
                                      1.6817928305074290860622509524664297901};
        if (fn > 0) width = w * kFwhmGaussian / fwhm[fn];
    }
    double weight(double y) const {
        const double z = (y - opt) / width;
        switch (fn) {
        case ORC_EXP_GAUSSIAN: return exp(-0.5 * z * z);
        case ORC_EXP_LORENTZIAN: return 1.0 / (1.0 + 0.5 * z * z);
        case ORC_EXP_HALFSINE: return fabs(z) <= M_PI_2 ? cos(z) : 0.0;
        case ORC_EXP_FULLSINE: return fabs(z) <= M_PI ? 0.5 + cos(z) / 2.0 : 0.0;
        default:
            if (fabs(z) <= 1.0) { const double z2 = z * z; return (1.0 - z2) * (1.0 + z2); }
            return 0.0;
        }
    }
};

// y = gray / max: double for integer images, float for float images
// (src/enfuse.h:637).
template <class T> double exposure_y(T gray);
template <> double exposure_y<uint8_t>(uint8_t g) { return (double)g / 255.0; }
template <> double exposure_y<uint16_t>(uint16_t g) { return (double)g / 65535.0; }
template <> double exposure_y<float>(float g) { return (double)(g / FLT_MAX); }

template <class T> struct GrayIndex { static int of(T g) { return (int)g; } };

// ---- local standard deviation, src/enfuse.h:75-219 -----------------------
// S = RealPromote of the gray type (double for int, float for float).
// FRESH = true is a deliberately wrong variant that sums every window anew
// instead of sliding the running sums (reachable only through
// orc_local_stddev_f32, so that tests can show they detect it).
template <class G, class S, bool FRESH = false>
void local_stddev(const G* gray, const uint8_t* mask, int w, int h, int n, G* out)
{
    const int b = n / 2;
#pragma omp parallel for schedule(static)
    for (int row = 0; row < h - 2 * b; ++row) {
        const int yc = row + b;
        std::vector<S> cs(w + 1), cq(w + 1);
        std::vector<size_t> cn(w + 1);
        S sum = S(0), sumSqr = S(0);
        size_t cnt = 0;
        auto column = [&](int x, S& s, S& q, size_t& k) {
            s = S(0); q = S(0); k = 0;
            for (int y = yc - b; y <= yc + b; ++y) {
                if (mask[(long)y * w + x]) {
                    const S v = (S)gray[(long)y * w + x];
                    s += v;
                    q += v * v;
                    ++k;
                }
            }
        };
        for (int x = 0; x <= 2 * b; ++x) {
            column(x, cs[x], cq[x], cn[x]);
            sum += cs[x];
            sumSqr += cq[x];
            cnt += cn[x];
        }
        for (int xc = b;; ++xc) {
            if (mask[(long)yc * w + xc]) {
                const S r = cnt <= 1 ? S(0) : (S)sqrt((sumSqr - sum * sum / (S)cnt) / (S)(cnt - 1));
                out[(long)yc * w + xc] = fromRealPromote<G>((double)r);
            }
            if (xc == w - b - 1) break;
            const int xn = xc + b + 1;
            S s, q;
            size_t k;
            column(xn, s, q, k);
            cs[xn] = s; cq[xn] = q; cn[xn] = k;
            if (FRESH) {
                sum = S(0); sumSqr = S(0); cnt = 0;
                for (int x = xn - n + 1; x <= xn; ++x) { sum += cs[x]; sumSqr += cq[x]; cnt += cn[x]; }
                continue;
            }
            sum += s - cs[xn - n];
            sumSqr += q - cq[xn - n];
            cnt += k - cn[xn - n];
        }
    }
}

// ---- local entropy, src/enfuse.h:222-566 ---------------------------------
// Exact-value histograms per channel; the entropy sums p*log(p) over the bins
// in ascending key order, using precomputed tables when the window is full.
template <class T>
struct Hist {
    std::map<T, unsigned> h[3];
    unsigned total[3] = {0, 0, 0};
    void clear() { for (int c = 0; c < 3; ++c) { h[c].clear(); total[c] = 0; } }
    void insert(int c, T k, unsigned n) {
        auto it = h[c].lower_bound(k);
        if (it != h[c].end() && !(k < it->first)) it->second += n;
        else h[c].insert(it, std::make_pair(k, n));
        total[c] += n;
    }
    void erase(int c, T k, unsigned n) {
        auto it = h[c].find(k);
        if (it->second > n) it->second -= n;
        else h[c].erase(it);
        total[c] -= n;
    }
    void insert(const Hist& o) { for (int c = 0; c < 3; ++c) for (auto& kv : o.h[c]) insert(c, kv.first, kv.second); }
    void erase(const Hist& o) { for (int c = 0; c < 3; ++c) for (auto& kv : o.h[c]) erase(c, kv.first, kv.second); }
};

template <class T>
void local_entropy(const T* img, const uint8_t* mask, int w, int h, int n, T* out)
{
    const size_t full = (size_t)n * n;
    std::vector<double> plog(full + 1), pent(full + 1);
    plog[0] = 0.0; pent[0] = 0.0;
    for (size_t i = 1; i <= full; ++i) {
        const double p = (double)i / (double)full;
        plog[i] = log((double)i);
        pent[i] = p * log(p);
    }
    const double maxv = Max<T>::v();
    auto entropy_of = [&](const Hist<T>& H, int c) -> double {
        const unsigned total = H.total[c];
        const size_t bins = H.h[c].size();
        if (total == 0 || bins <= 1) return 0.0;
        double e = 0.0;
        if (total == full) {
            for (auto& kv : H.h[c]) e += pent[kv.second];
            return -e / plog[bins];
        }
        for (auto& kv : H.h[c]) {
            const double p = kv.second / (double)total;
            e += p * log(p);
        }
        return -e / log((double)bins);
    };
    const int b = n / 2;
    std::vector<Hist<T>> pad(h + 1);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < n; ++x)
            if (mask[(long)y * w + x])
                for (int c = 0; c < 3; ++c) pad[y].insert(c, img[((long)y * w + x) * 3 + c], 1);
    Hist<T> hist;
    for (int xc = b; xc < w - b; ++xc) {
        hist.clear();
        for (int y = 0; y <= 2 * b; ++y) hist.insert(pad[y]);
        for (int yc = b; yc < h - b; ++yc) {
            if (mask[(long)yc * w + xc])
                for (int c = 0; c < 3; ++c) out[((long)yc * w + xc) * 3 + c] = fromRealPromote<T>(entropy_of(hist, c) * maxv);
            hist.erase(pad[yc - b]);
            hist.insert(pad[yc + b + 1]);
        }
        for (int y = 0; y < h; ++y) {
            if (mask[(long)y * w + xc - b])
                for (int c = 0; c < 3; ++c) pad[y].erase(c, img[((long)y * w + xc - b) * 3 + c], 1);
            if (xc + b + 1 < w && mask[(long)y * w + xc + b + 1])
                for (int c = 0; c < 3; ++c) pad[y].insert(c, img[((long)y * w + xc + b + 1) * 3 + c], 1);
        }
    }
}

template <class T> struct Contrast;   // gray type used by the contrast weight
template <> struct Contrast<uint8_t> { typedef int32_t G; typedef double S; };
template <> struct Contrast<uint16_t> { typedef int32_t G; typedef double S; };
template <> struct Contrast<float> { typedef float G; typedef float S; };

template <class T> float contrast_weight(double wc, typename Contrast<T>::G g);
template <> float contrast_weight<uint8_t>(double wc, int32_t g) { return (float)(wc * (double)g / 255.0); }
template <> float contrast_weight<uint16_t>(double wc, int32_t g) { return (float)(wc * (double)g / 65535.0); }
template <> float contrast_weight<float>(double wc, float g) { return (float)(wc * (double)g); }

// ---- Laplacian-of-Gaussian contrast, src/enfuse.h:1020-1093 --------------
// VIGRA's pieces (not vendored; restated from its documented algorithms):
// Kernel1D<double>::initGaussian / initGaussianDerivative (radius
// (int)(3 sigma + 0.5), resp. (int)((3 + order / 2) sigma + 0.5), at least 1;
// samples of the Gaussian functor, derivative kernels with their mean
// removed, normalised so that sum k = 1, resp. sum k (-x)^2 / 2! = 1),
// convolveLine with BORDER_TREATMENT_REFLECT (sum over the window from the
// leftmost source sample, source index mirrored at both ends without
// repeating the edge, "kernel longer than line" when radius >= length),
// separableConvolveX/Y, laplacianOfGaussian (d2x * smooth_y + smooth_x *
// d2y), gaussianSmoothing, gaussianSharpening ((1 + f) src - f smooth).
std::vector<double> vigra_gaussian_kernel(double sigma, int order)
{
    int radius = order == 0 ? (int)(3.0 * sigma + 0.5) : (int)((3.0 + 0.5 * order) * sigma + 0.5);
    if (radius == 0) radius = 1;
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

inline int reflect(int i, int n) { return i < 0 ? -i : (i >= n ? 2 * (n - 1) - i : i); }

// One separable pass: out(x, y) = cast(sum over t of k[t] * in(x - r + t, y))
// along rows (dir 0) or columns (dir 1).
template <class I, class O>
void convolve_pass(const I* in, int w, int h, const std::vector<double>& k, int dir, O* out)
{
    const int r = (int)(k.size() / 2);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            double sum = 0.0;
            for (int t = 0; t <= 2 * r; ++t) {
                const I v = dir == 0 ? in[(long)y * w + reflect(x - r + t, w)] : in[(long)reflect(y - r + t, h) * w + x];
                sum += k[(size_t)t] * (double)v;
            }
            out[(long)y * w + x] = fromRealPromote<O>(sum);
        }
}

template <class T> struct LoGTypes;   // gray type G and the convolution images' type R (RealPromote<G>)
template <> struct LoGTypes<uint8_t> { typedef int32_t G; typedef double R; };
template <> struct LoGTypes<uint16_t> { typedef int32_t G; typedef double R; };
template <> struct LoGTypes<float> { typedef float G; typedef float R; };

template <class G> G max_of();
template <> int32_t max_of<int32_t>() { return INT32_MAX; }
template <> float max_of<float>() { return FLT_MAX; }

// grad = the contrast image of the edge mode, written where alpha is set
// (zero elsewhere); lc = the local standard deviation when min_curvature > 0.
template <class T>
int log_contrast(const typename LoGTypes<T>::G* gray, const uint8_t* alpha, int w, int h,
                 const orc_enfuse_params* p, const typename LoGTypes<T>::G* lc, typename LoGTypes<T>::G* grad)
{
    typedef typename LoGTypes<T>::G G;
    typedef typename LoGTypes<T>::R R;
    const long N = (long)w * h;
    const std::vector<double> smooth = vigra_gaussian_kernel(p->edge_scale, 0);
    const std::vector<double> deriv = vigra_gaussian_kernel(p->edge_scale, 2);
    const int rs = (int)(smooth.size() / 2), rd = (int)(deriv.size() / 2);
    if (rs >= w || rs >= h || rd >= w || rd >= h) return -1;   // convolveLine(): kernel longer than line
    std::vector<G> lce;
    const G* src = gray;
    if (p->lce_scale > 0.0) {
        const std::vector<double> ls = vigra_gaussian_kernel(p->lce_scale, 0);
        const int rl = (int)(ls.size() / 2);
        if (rl >= w || rl >= h) return -1;
        std::vector<R> t1((size_t)N), t2((size_t)N);
        convolve_pass<G, R>(gray, w, h, ls, 0, t1.data());
        convolve_pass<R, R>(t1.data(), w, h, ls, 1, t2.data());
        lce.resize((size_t)N);
        const double f = p->lce_factor;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < N; ++i) lce[i] = fromRealPromote<G>((1.0 + f) * (double)gray[i] - f * (double)t2[i]);
        src = lce.data();
    }
    std::vector<R> tmp((size_t)N), tx((size_t)N), ty((size_t)N);
    convolve_pass<G, R>(src, w, h, deriv, 0, tmp.data());
    convolve_pass<R, R>(tmp.data(), w, h, smooth, 1, tx.data());
    convolve_pass<G, R>(src, w, h, smooth, 0, tmp.data());
    convolve_pass<R, R>(tmp.data(), w, h, deriv, 1, ty.data());
    const double mc = p->min_curvature;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
        const R sum = (R)(tx[i] + ty[i]);
        const G lap = fromRealPromote<G>((double)std::fabs(sum));   // MagnitudeAccessor
        if (!alpha[i]) { grad[i] = G(0); continue; }
        if (mc <= 0.0) {
            // ClampingFunctor(-mc -> 0, max -> max)
            const G lo = (G)(-mc), up = max_of<G>();
            grad[i] = lap <= lo ? G(0) : (lap >= up ? up : lap);
        } else {
            // FillInFunctor(mc, 1, mc / max(ScalarType))
            const G thr = (G)mc;
            const double s2 = mc / Max<T>::v();
            grad[i] = lap >= thr ? fromRealPromote<G>(1.0 * (double)lap) : fromRealPromote<G>(s2 * (double)lc[i]);
        }
    }
    return 0;
}

template <class T> bool is_integral() { return true; }
template <> bool is_integral<float>() { return false; }

// enfuseMask, src/enfuse.h:962-1212.
template <class T>
int enfuse_mask_t(const T* img, const uint8_t* alpha, int w, int h, const orc_enfuse_params* p, float* mask)
{
    const long N = (long)w * h;
    for (long i = 0; i < N; ++i) mask[i] = 0.0f;
    if (p->gray_projector < ORC_GRAY_AVERAGE || p->gray_projector > ORC_GRAY_MIXER) return -1;
    if (p->gray_projector == ORC_GRAY_MIXER &&
        (p->mixer_red < 0.0 || p->mixer_green < 0.0 || p->mixer_blue < 0.0 ||
         p->mixer_red + p->mixer_green + p->mixer_blue == 0.0))
        return -1;   // check_weights (src/mga.h:150-180)

    if (p->w_exposure > 0.0) {
        const Exposure ex(p->exposure_function, p->exposure_optimum, p->exposure_width);
        // CutoffExposureFunctor (src/enfuse.h:654-720) with its default
        // projectors: the minimum channel ("anti-value") at or above the
        // lower cutoff and the maximum channel ("value") at or below the upper
        const bool cut = p->exposure_cutoff != 0;
        if (cut && (p->exposure_lower_cutoff < 0.0 || p->exposure_upper_cutoff < 0.0 ||
                    p->exposure_lower_cutoff > p->exposure_upper_cutoff))
            return -1;
#pragma omp parallel for schedule(guided)
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                const long i = (long)y * w + x;
                if (!alpha[i]) continue;
                const T* px = img + 3 * i;
                if (cut) {
                    const T lo = std::min(px[0], std::min(px[1], px[2]));
                    const T hi = std::max(px[0], std::max(px[1], px[2]));
                    if (!((double)lo >= p->exposure_lower_cutoff && (double)hi <= p->exposure_upper_cutoff)) {
                        mask[i] = 0.0f;
                        continue;
                    }
                }
                const T g = Gray<T, T>::get(px, p);
                double wt;
                if (p->exposure_table && is_integral<T>()) wt = p->exposure_table[GrayIndex<T>::of(g)];
                else wt = ex.weight(exposure_y<T>(g));
                mask[i] = (float)(p->w_exposure * wt);
            }
    }

    if (p->w_contrast > 0.0) {
        typedef typename Contrast<T>::G G;
        typedef typename Contrast<T>::S S;
        const int n = p->contrast_window;
        const bool edges = p->edge_scale > 0.0;
        const bool stddev = !edges || p->min_curvature > 0.0;
        if (stddev && (n < 2 || w < n || h < n)) return -1;
        std::vector<G> gray((size_t)N), grad((size_t)N, G(0));
#pragma omp parallel for schedule(guided)
        for (long i = 0; i < N; ++i) gray[i] = Gray<T, G>::get(img + 3 * i, p);
        if (!edges) {
            local_stddev<G, S>(gray.data(), alpha, w, h, n, grad.data());
        } else {
            std::vector<G> lc;
            if (stddev) {
                lc.assign((size_t)N, G(0));
                local_stddev<G, S>(gray.data(), alpha, w, h, n, lc.data());
            }
            if (p->lce_factor < 0.0 || p->lce_scale < 0.0) return -1;
            if (log_contrast<T>(gray.data(), alpha, w, h, p, lc.data(), grad.data())) return -1;
        }
#pragma omp parallel for schedule(guided)
        for (long i = 0; i < N; ++i)
            if (alpha[i]) mask[i] = contrast_weight<T>(p->w_contrast, grad[i]) + mask[i];
    }

    if (p->w_saturation > 0.0) {
        const double max_value = is_integral<T>() ? Max<T>::v() : 1.0;
#pragma omp parallel for schedule(guided)
        for (long i = 0; i < N; ++i) {
            if (!alpha[i]) continue;
            const T* q = img + 3 * i;
            const T mx = std::max(q[0], std::max(q[1], q[2]));
            const T mn = std::min(q[0], std::min(q[1], q[2]));
            float s = 0.0f;
            if (!(mx == mn)) {
                const double sum = (double)mx + (double)mn;
                const double diff = (double)mx - (double)mn;
                const double sat = sum <= max_value ? diff / sum : diff / (2.0 * max_value - sum);
                s = (float)(p->w_saturation * sat);
            }
            mask[i] = s + mask[i];
        }
    }

    if (p->w_entropy > 0.0) {
        const int n = p->entropy_window;
        if (w < n || h < n) return -1;
        std::vector<T> ent((size_t)N * 3, T(0));
        if (p->entropy_cutoff) {
            // the image clamped before the entropy (src/enfuse.h:1145-1188):
            // ClampingFunctor<RGB>(lower -> 0, upper -> max), whose blue
            // channel tests the red channel against the lower bound
            const double lc = p->entropy_lower_cutoff, uc = p->entropy_upper_cutoff;
            if (lc < 0.0 || uc < 0.0 || lc > uc) return -1;
            const T lo = (T)lc, up = (T)uc, mx = (T)Max<T>::v();
            std::vector<T> trunc((size_t)N * 3);
#pragma omp parallel for schedule(guided)
            for (long i = 0; i < N; ++i) {
                const T r = img[3 * i], g = img[3 * i + 1], b = img[3 * i + 2];
                trunc[3 * i] = r <= lo ? T(0) : (r >= up ? mx : r);
                trunc[3 * i + 1] = g <= lo ? T(0) : (g >= up ? mx : g);
                trunc[3 * i + 2] = r <= lo ? T(0) : (b >= up ? mx : b);
            }
            local_entropy<T>(trunc.data(), alpha, w, h, n, ent.data());
        } else {
            local_entropy<T>(img, alpha, w, h, n, ent.data());
        }
#pragma omp parallel for schedule(guided)
        for (long i = 0; i < N; ++i) {
            if (!alpha[i]) continue;
            const T mn = std::min(std::min(ent[3 * i], ent[3 * i + 1]), ent[3 * i + 2]);
            mask[i] = (float)(p->w_entropy * (double)(float)mn) + mask[i];
        }
    }
    return 0;
}

template <class P> int ptype_of();
template <> int ptype_of<int16_t>() { return ORC_I16; }
template <> int ptype_of<int32_t>() { return ORC_I32; }
template <> int ptype_of<double>() { return ORC_F64; }

// enfuseMain, src/enfuse.h:1217-1609 (soft or hard masks).
template <class T, class P, class M>
int enfuse_t(int w, int h, int K, const T* const* imgs, const uint8_t* const* alphas,
             const orc_enfuse_params* prm, int levels, bool wrap, int img_frac, int mask_frac,
             T* out, uint8_t* out_alpha, int dither, uint64_t seed)
{
    const long N = (long)w * h;
    std::vector<std::vector<float>> masks(K, std::vector<float>((size_t)N));
    std::vector<float> norm((size_t)N, 0.0f);
    std::vector<uint8_t> ua((size_t)N, 0);
    for (int m = 0; m < K; ++m) {
        int rc = enfuse_mask_t<T>(imgs[m], alphas[m], w, h, prm, masks[m].data());
        if (rc) return rc;
#pragma omp parallel for schedule(guided)
        for (long i = 0; i < N; ++i) {
            if (alphas[m][i]) ua[i] = alphas[m][i];
            norm[i] = masks[m][i] + norm[i];
        }
    }
    if (prm->hard_mask) {
        // Hard masks (src/enfuse.h:1402-1443): per pixel the first image of
        // largest weight above 0 gets 255, the others 0; where no weight is
        // above 0 every image gets float(255) / K.
#pragma omp parallel for schedule(guided)
        for (long i = 0; i < N; ++i) {
            float mx = 0.0f;
            int maxi = 0;
            for (int m = 0; m < K; ++m) {
                const float wv = masks[m][i];
                if (wv > mx) { mx = wv; maxi = m; }
            }
            for (int m = 0; m < K; ++m)
                masks[m][i] = mx == 0.0f ? (float)255 / K : (m == maxi ? 255.0f : 0.0f);
        }
    }
    long nimg = 0, nmask = 0;
    std::vector<Dims> di = pyramid_dims(w, h, levels, 3, &nimg);
    std::vector<Dims> dm = pyramid_dims(w, h, levels, 1, &nmask);
    std::vector<P> result;
    const float fallback = (float)(255 / K);
    const double max_mask_pyr = (double)((M)(255 << mask_frac));
    const double scale = (double)(1u << mask_frac);
    // A CIELAB / CIELUV blend space: the pyramids are built from the images
    // converted to the scaled blend space, the result goes back to RGB under
    // the union alpha (src/fixmath.h:2018-2294); the weights above used RGB.
    // (CIECAM02 enters and leaves the pyramids through image_to_level0 /
    // level0_to_image themselves.)
    const int cs_all = std::is_same<T, float>::value ? ORC_COLORSPACE_IDENTITY : orc_blend_colorspace();
    const int cs = cs_all == ORC_COLORSPACE_CIECAM ? ORC_COLORSPACE_IDENTITY : cs_all;
    const int cstype = sizeof(T) == 1 ? ORC_U8 : ORC_U16;
    std::vector<T> scaled(cs != ORC_COLORSPACE_IDENTITY ? (size_t)N * 3 : 0);
    for (int m = 0; m < K; ++m) {
        std::vector<P> lp((size_t)nimg);
        const T* src = imgs[m];
        if (cs != ORC_COLORSPACE_IDENTITY) {
            orc_rgb_to_blend_space(cs, cstype, w, h, imgs[m], scaled.data());
            src = scaled.data();
        }
        image_to_level0(img_frac, 3, src, w, h, lp.data());
        int rc = gaussian_levels<P>(lp.data(), di, 3, wrap, alphas[m]);
        if (rc) return rc;
        laplacian_levels<P>(lp.data(), di, 3, wrap);

        // Normalise and convert to fixed point (src/enfuse.h:1529-1546,
        // src/fixmath.h:188-194).
        std::vector<M> mgp((size_t)nmask);
        const std::vector<float>& mk = masks[m];
#pragma omp parallel for schedule(guided)
        for (long i = 0; i < N; ++i) {
            const float nm = prm->hard_mask ? mk[i] : norm[i] > 0.0 ? (255.0f * mk[i]) / norm[i] : fallback;
            mgp[i] = fromRealPromote<M>((double)nm * scale);
        }
        rc = gaussian_levels<M>(mgp.data(), dm, 1, wrap, ua.data());
        if (rc) return rc;

        // ImageMaskMultiplyFunctor (src/enfuse.h:601-620) and accumulation.
        for (int l = 0; l < levels; ++l) {
            const long n = (long)di[l].w * di[l].h;
            P* L = lp.data() + di[l].off;
            const M* G = mgp.data() + dm[l].off;
#pragma omp parallel for schedule(guided)
            for (long i = 0; i < n; ++i) {
                const double coeff = (double)G[i] / max_mask_pyr;
                for (int c = 0; c < 3; ++c) L[c * n + i] = fromRealPromote<P>((double)L[c * n + i] * coeff);
            }
        }
        if (m == 0) {
            result.swap(lp);
        } else {
#pragma omp parallel for schedule(guided)
            for (long i = 0; i < nimg; ++i) result[i] = Store<P>::wrap(Store<P>::A::add((typename Store<P>::A::T)lp[i], (typename Store<P>::A::T)result[i]));
        }
    }
    collapse_levels<P>(result.data(), di, 3, wrap);
    if (cs != ORC_COLORSPACE_IDENTITY) {
        level0_to_image(img_frac, 3, result.data(), w, h, ua.data(), scaled.data(), dither, seed);
        orc_blend_space_to_rgb_if(cs, cstype, w, h, scaled.data(), ua.data(), out);
    } else {
        level0_to_image(img_frac, 3, result.data(), w, h, ua.data(), out, dither, seed);
    }
    memcpy(out_alpha, ua.data(), (size_t)N);
    return 0;
}

}  // namespace

extern "C" {

// The float local standard deviation of the contrast weight on its own
// (localStdDevIf over a float gray image, RealPromote<float> = float sums).
// planted 1 and 2 are deliberately wrong variants for tests that must
// detect them: 1 accumulates in double, 2 sums every window anew.
int orc_local_stddev_f32(const float* gray, const uint8_t* mask, int w, int h, int n, int planted, float* out)
{
    if (n < 3 || !(n & 1) || w < n || h < n || planted < 0 || planted > 2) return -1;
    if (planted == 1) local_stddev<float, double>(gray, mask, w, h, n, out);
    else if (planted == 2) local_stddev<float, float, true>(gray, mask, w, h, n, out);
    else local_stddev<float, float>(gray, mask, w, h, n, out);
    return 0;
}

int orc_enfuse_mask(int itype, const void* img, const uint8_t* alpha, int w, int h,
                    const orc_enfuse_params* prm, float* mask)
{
    switch (itype) {
    case ORC_U8: return enfuse_mask_t<uint8_t>((const uint8_t*)img, alpha, w, h, prm, mask);
    case ORC_U16: return enfuse_mask_t<uint16_t>((const uint16_t*)img, alpha, w, h, prm, mask);
    case ORC_F32: return enfuse_mask_t<float>((const float*)img, alpha, w, h, prm, mask);
    }
    return -4;
}

int orc_enfuse(int itype, int w, int h, int K, const void* const* imgs, const uint8_t* const* alphas,
               const orc_enfuse_params* prm, int levels, int wrap, void* out, uint8_t* out_alpha,
               int dither, uint64_t seed)
{
    if (K < 1 || levels < 1 || levels > 29) return -1;
    switch (itype) {
    case ORC_U8:
        return enfuse_t<uint8_t, int16_t, int16_t>(w, h, K, (const uint8_t* const*)imgs, alphas, prm, levels, wrap, 7, 7,
                                                   (uint8_t*)out, out_alpha, dither, seed);
    case ORC_U16:
        return enfuse_t<uint16_t, int32_t, int32_t>(w, h, K, (const uint16_t* const*)imgs, alphas, prm, levels, wrap, 7, 15,
                                                    (uint16_t*)out, out_alpha, dither, seed);
    case ORC_F32:
        return enfuse_t<float, double, int32_t>(w, h, K, (const float* const*)imgs, alphas, prm, levels, wrap, 0, 15,
                                                (float*)out, out_alpha, dither, seed);
    }
    return -4;
}

}  // extern "C"

// VIGRA's Gaussian (order 0) or second-derivative (order 2) kernel of scale
// sigma, 2r + 1 taps into k (capacity n); returns r, or -1 when n is short.
extern "C" int orc_vigra_kernel(double sigma, int order, double* k, int n)
{
    const std::vector<double> v = vigra_gaussian_kernel(sigma, order);
    if ((int)v.size() > n) return -1;
    for (size_t i = 0; i < v.size(); ++i) k[i] = v[i];
    return (int)(v.size() / 2);
}
