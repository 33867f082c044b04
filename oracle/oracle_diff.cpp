// FP64 CPU restatement of the relaxed differentiable enfuse, forward and
// adjoint (TEST INFRASTRUCTURE ONLY: checker of the HIP adjoint path).
//
// The relaxed variant is not the reference's arithmetic: the reference's
// fixed-point mask pyramid has zero gradient almost everywhere and its float
// exposure weight is constant (y = gray / FLT_MAX, src/enfuse.h:637), so the
// differentiable form keeps every quantity in double:
//   gray   = (r + g + b) / 3                                (src/mga.h:188-193)
//   E      = exp(-((gray - mu) / sigma)^2 / 2)              (Gaussian, src/exposure_weight.h:44-60)
//   s      = (max - min) / (max + min + 1e-6)               (smooth form of src/enfuse.h:733-780)
//   sd     = sqrt(sum (q - mean)^2 / (N - 1) + 1e-12) over the n x n window, interior
//            pixels only, 0 at the borders                  (src/enfuse.h:75-219)
//   m      = w_e E + w_c sd + w_s s + 1e-12                 (order of src/enfuse.h:1119-1135)
//   m^     = m / sum_k m                                    (src/enfuse.h:1526-1535)
//   G_k    = Gaussian pyramid of T(I_k), T(v) = v >= 0 ? 1 + log1p(v) : 1 / (1 - v)
//   M_k    = Gaussian pyramid of m^_k                       (reduce with an all-set alpha)
//   R_l    = sum_k (G_kl - expand(G_k,l+1)) M_kl            (Laplacian on the fly, merge)
//   C_l    = R_l + expand(C_l+1), O = T^-1(C_0)             (collapse, src/fixmath.h:320-329)
// The pyramid operations are the reference's double reduce/expand
// (oracle_core.h).  The adjoint runs the transposed operations:
//   expand^T: gather of g / (hsum * vsum) at stride 2 with the expand taps
//             (columns 1 4 6 4 1, rows 1 1 6 1 1),
//   reduce^T: polyphase [1 6 1] / [4 4] scatter of g / (in-range weight sum).
#include <math.h>
#include <string.h>

#include <vector>

#include "oracle.h"
#include "oracle_core.h"

namespace {

using orc::Dims;

constexpr double kEpsSat = 1e-6;
constexpr double kEpsSd = 1e-12;
constexpr double kEpsW = 1e-12;
constexpr int kBin[5] = {1, 4, 6, 4, 1};

struct Params { double we, ws, wc, mu, sigma; };

double T(double v) { return v >= 0.0 ? 1.0 + log1p(v) : 1.0 / (1.0 - v); }
double dT(double v) { return v >= 0.0 ? 1.0 / (1.0 + v) : 1.0 / ((1.0 - v) * (1.0 - v)); }
double Tinv(double c) { return c >= 1.0 ? expm1(c - 1.0) : 1.0 - 1.0 / c; }
double dTinv(double c) { return c >= 1.0 ? exp(c - 1.0) : 1.0 / (c * c); }

// Forward per-pixel weight terms of one image.
struct Terms { std::vector<double> gray, E, s, sd, mean; std::vector<int> imx, imn; };

void weight_terms(const double* img, int w, int h, int n, const Params& p, Terms& t)
{
    const long N = (long)w * h;
    t.gray.assign(N, 0.0); t.E.assign(N, 0.0); t.s.assign(N, 0.0); t.sd.assign(N, 0.0); t.mean.assign(N, 0.0);
    t.imx.assign(N, 0); t.imn.assign(N, 0);
    for (long i = 0; i < N; ++i) {
        const double* q = img + 3 * i;
        t.gray[i] = (q[0] + q[1] + q[2]) / 3.0;
        const double z = (t.gray[i] - p.mu) / p.sigma;
        t.E[i] = exp(-0.5 * z * z);
        int a = 0, b = 0;
        for (int c = 1; c < 3; ++c) {
            if (q[c] > q[a]) a = c;
            if (q[c] < q[b]) b = c;
        }
        t.imx[i] = a;
        t.imn[i] = b;
        t.s[i] = (q[a] - q[b]) / (q[a] + q[b] + kEpsSat);
    }
    const int r = n / 2;
    const double Nw = (double)n * n;
#pragma omp parallel for
    for (int y = r; y < h - r; ++y)
        for (int x = r; x < w - r; ++x) {
            double sum = 0.0;
            for (int yy = y - r; yy <= y + r; ++yy)
                for (int xx = x - r; xx <= x + r; ++xx) sum += t.gray[(long)yy * w + xx];
            const double mean = sum / Nw;
            double ss = 0.0;
            for (int yy = y - r; yy <= y + r; ++yy)
                for (int xx = x - r; xx <= x + r; ++xx) {
                    const double d = t.gray[(long)yy * w + xx] - mean;
                    ss += d * d;
                }
            t.mean[(long)y * w + x] = mean;
            t.sd[(long)y * w + x] = sqrt(ss / (Nw - 1.0) + kEpsSd);
        }
}

// expand^T: coarse gradient (cw x ch) += transpose of expand applied to the
// fine gradient g (fw x fh).
void expand_T(const double* g, int fw, int fh, double* gc, int cw, int chh)
{
    for (int Y = 0; Y < fh; ++Y)
        for (int X = 0; X < fw; ++X) {
            const double v = g[(long)Y * fw + X] / (double)(orc::expand_hsum(X, cw, false, (fw & 1) == 0) * orc::expand_vsum(Y, chh));
            const int j = X >> 1, i = Y >> 1;
            int jx[3], wx[3], nx = 0, iy[3], wy[3], ny = 0;
            if ((X & 1) == 0) {
                if (j - 1 >= 0) { jx[nx] = j - 1; wx[nx++] = 1; }
                jx[nx] = j; wx[nx++] = 6;
                if (j + 1 < cw) { jx[nx] = j + 1; wx[nx++] = 1; }
            } else {
                jx[nx] = j; wx[nx++] = 4;
                if (j + 1 < cw) { jx[nx] = j + 1; wx[nx++] = 4; }
            }
            if ((Y & 1) == 0) {
                if (i - 1 >= 0) { iy[ny] = i - 1; wy[ny++] = 1; }
                iy[ny] = i; wy[ny++] = 6;
                if (i + 1 < chh) { iy[ny] = i + 1; wy[ny++] = 1; }
            } else {
                // odd rows weigh rows i and i+1 equally: [1 1] over a sum of 1 or 2
                // (expand_vsum), while odd columns use [4 4] over 4 or 8 (expand_hsum)
                iy[ny] = i; wy[ny++] = 1;
                if (i + 1 < chh) { iy[ny] = i + 1; wy[ny++] = 1; }
            }
            for (int a = 0; a < ny; ++a)
                for (int b = 0; b < nx; ++b) gc[(long)iy[a] * cw + jx[b]] += (double)(wy[a] * wx[b]) * v;
        }
}

// reduce^T (all-set alpha): fine gradient (sw x sh) += transpose of reduce
// applied to the coarse gradient g (dw x dh).
void reduce_T(const double* g, int dw, int dh, double* gf, int sw, int sh)
{
    for (int i = 0; i < dh; ++i)
        for (int j = 0; j < dw; ++j) {
            int wxs = 0, wys = 0;
            for (int d = -2; d <= 2; ++d) {
                if (2 * j + d >= 0 && 2 * j + d < sw) wxs += kBin[d + 2];
                if (2 * i + d >= 0 && 2 * i + d < sh) wys += kBin[d + 2];
            }
            const double v = g[(long)i * dw + j] / (double)(wxs * wys);
            for (int dy = -2; dy <= 2; ++dy) {
                const int y = 2 * i + dy;
                if (y < 0 || y >= sh) continue;
                for (int dx = -2; dx <= 2; ++dx) {
                    const int x = 2 * j + dx;
                    if (x < 0 || x >= sw) continue;
                    gf[(long)y * sw + x] += (double)(kBin[dy + 2] * kBin[dx + 2]) * v;
                }
            }
        }
}

struct Forward {
    int w, h, K, L;
    std::vector<Dims> d3, d1;
    long tot3 = 0, tot1 = 0;
    std::vector<std::vector<double>> G;   // K image Gaussian pyramids (3 planes per level)
    std::vector<std::vector<double>> M;   // K mask Gaussian pyramids
    std::vector<Terms> terms;
    std::vector<double> m, S;             // weights (K planes), norm
    std::vector<double> C;                // collapsed result pyramid (3 planes per level)
};

int forward(int w, int h, int K, const double* imgs, const Params& p, int L, int n, Forward& F)
{
    F.w = w; F.h = h; F.K = K; F.L = L;
    F.d3 = orc::pyramid_dims(w, h, L, 3, &F.tot3);
    F.d1 = orc::pyramid_dims(w, h, L, 1, &F.tot1);
    const long N = (long)w * h;
    std::vector<uint8_t> ones(N, 255);
    F.terms.resize(K);
    F.m.assign((size_t)K * N, 0.0);
    F.S.assign(N, 0.0);
    for (int k = 0; k < K; ++k) {
        weight_terms(imgs + (size_t)k * N * 3, w, h, n, p, F.terms[k]);
        const Terms& t = F.terms[k];
        for (long i = 0; i < N; ++i) {
            double v = p.we * t.E[i];
            v = v + p.wc * t.sd[i];
            v = v + p.ws * t.s[i];
            v = v + kEpsW;
            F.m[(size_t)k * N + i] = v;
            F.S[i] = F.S[i] + v;
        }
    }
    F.G.assign(K, std::vector<double>(F.tot3));
    F.M.assign(K, std::vector<double>(F.tot1));
    for (int k = 0; k < K; ++k) {
        const double* im = imgs + (size_t)k * N * 3;
        for (int c = 0; c < 3; ++c)
            for (long i = 0; i < N; ++i) F.G[k][(long)c * N + i] = T(im[3 * i + c]);
        for (long i = 0; i < N; ++i) F.M[k][i] = F.m[(size_t)k * N + i] / F.S[i];
        int rc = orc::gaussian_levels<double>(F.G[k].data(), F.d3, 3, false, ones.data());
        if (!rc) rc = orc::gaussian_levels<double>(F.M[k].data(), F.d1, 1, false, ones.data());
        if (rc) return rc;
    }
    // Merge with Laplacians evaluated per level, then collapse.
    F.C.assign(F.tot3, 0.0);
    for (int l = 0; l < L; ++l) {
        const Dims& a = F.d3[l];
        const long n_l = (long)a.w * a.h;
        for (int k = 0; k < K; ++k) {
            std::vector<double> lap(F.G[k].begin() + a.off, F.G[k].begin() + a.off + 3 * n_l);
            if (l + 1 < L) {
                const Dims& b = F.d3[l + 1];
                for (int c = 0; c < 3; ++c)
                    orc::expand_t<double>(false, false, F.G[k].data() + b.off + (long)c * b.w * b.h, b.w, b.h, b.w,
                                          lap.data() + (long)c * n_l, a.w, a.h, a.w);
            }
            const double* mk = F.M[k].data() + F.d1[l].off;
            for (int c = 0; c < 3; ++c)
                for (long i = 0; i < n_l; ++i) {
                    double& r = F.C[a.off + (long)c * n_l + i];
                    r = r + lap[(long)c * n_l + i] * mk[i];
                }
        }
    }
    orc::collapse_levels<double>(F.C.data(), F.d3, 3, false);
    return 0;
}

}  // namespace

extern "C" {

int orc_diff_forward(int w, int h, int K, const double* imgs, const double* prm, int levels, int n, double* out)
{
    if (w < 1 || h < 1 || K < 1 || levels < 1 || n < 3 || !(n & 1)) return -3;
    const Params p{prm[0], prm[1], prm[2], prm[3], prm[4]};
    Forward F;
    int rc = forward(w, h, K, imgs, p, levels, n, F);
    if (rc) return rc;
    const long N = (long)w * h;
    for (long i = 0; i < N; ++i)
        for (int c = 0; c < 3; ++c) out[3 * i + c] = Tinv(F.C[(long)c * N + i]);
    return 0;
}

// The transposes alone, for their dot-product tests: g_coarse += expand^T(g_fine),
// g_fine += reduce^T(g_coarse) (all-set alpha).
void orc_diff_expand_T(const double* g, int fw, int fh, double* gc, int cw, int ch) { expand_T(g, fw, fh, gc, cw, ch); }
void orc_diff_reduce_T(const double* g, int dw, int dh, double* gf, int sw, int sh) { reduce_T(g, dw, dh, gf, sw, sh); }

int orc_diff_backward(int w, int h, int K, const double* imgs, const double* prm, int levels, int n,
                      const double* gout, double* gimgs, double* gprm)
{
    if (w < 1 || h < 1 || K < 1 || levels < 1 || n < 3 || !(n & 1)) return -3;
    const Params p{prm[0], prm[1], prm[2], prm[3], prm[4]};
    Forward F;
    int rc = forward(w, h, K, imgs, p, levels, n, F);
    if (rc) return rc;
    const int L = levels;
    const long N = (long)w * h;
    // Adjoint of the output transform and of the collapse: gC_l, finest first.
    std::vector<double> gC(F.tot3, 0.0);
    for (long i = 0; i < N; ++i)
        for (int c = 0; c < 3; ++c) gC[(long)c * N + i] = gout[3 * i + c] * dTinv(F.C[(long)c * N + i]);
    for (int l = 0; l + 1 < L; ++l) {
        const Dims& a = F.d3[l];
        const Dims& b = F.d3[l + 1];
        for (int c = 0; c < 3; ++c)
            expand_T(gC.data() + a.off + (long)c * a.w * a.h, a.w, a.h, gC.data() + b.off + (long)c * b.w * b.h, b.w, b.h);
    }
    // Per image: merge adjoint, Laplacian + Gaussian adjoint, mask pyramid adjoint.
    std::vector<double> gmh((size_t)K * N, 0.0);   // d loss / d m^_k
    for (int k = 0; k < K; ++k) {
        std::vector<double> gL(F.tot3, 0.0), gM(F.tot1, 0.0);
        for (int l = 0; l < L; ++l) {
            const Dims& a = F.d3[l];
            const long n_l = (long)a.w * a.h;
            std::vector<double> lap(F.G[k].begin() + a.off, F.G[k].begin() + a.off + 3 * n_l);
            if (l + 1 < L) {
                const Dims& b = F.d3[l + 1];
                for (int c = 0; c < 3; ++c)
                    orc::expand_t<double>(false, false, F.G[k].data() + b.off + (long)c * b.w * b.h, b.w, b.h, b.w,
                                          lap.data() + (long)c * n_l, a.w, a.h, a.w);
            }
            const double* mk = F.M[k].data() + F.d1[l].off;
            for (long i = 0; i < n_l; ++i) {
                double gm = 0.0;
                for (int c = 0; c < 3; ++c) {
                    const double g = gC[a.off + (long)c * n_l + i];
                    gL[a.off + (long)c * n_l + i] = g * mk[i];
                    gm += g * lap[(long)c * n_l + i];
                }
                gM[F.d1[l].off + i] = gm;
            }
        }
        // Image pyramid: gG_l = gL_l - expand^T(gL_l-1) + reduce^T(gG_l+1), coarsest first.
        std::vector<double> gG(F.tot3, 0.0);
        for (int l = L - 1; l >= 0; --l) {
            const Dims& a = F.d3[l];
            const long n_l = (long)a.w * a.h;
            for (long i = 0; i < 3 * n_l; ++i) gG[a.off + i] = gL[a.off + i];
            if (l >= 1) {
                const Dims& f = F.d3[l - 1];
                std::vector<double> t(3 * n_l, 0.0);
                for (int c = 0; c < 3; ++c)
                    expand_T(gL.data() + f.off + (long)c * f.w * f.h, f.w, f.h, t.data() + (long)c * n_l, a.w, a.h);
                for (long i = 0; i < 3 * n_l; ++i) gG[a.off + i] -= t[i];
            }
            if (l + 1 < L) {
                const Dims& b = F.d3[l + 1];
                for (int c = 0; c < 3; ++c)
                    reduce_T(gG.data() + b.off + (long)c * b.w * b.h, b.w, b.h, gG.data() + a.off + (long)c * n_l, a.w, a.h);
            }
        }
        const double* im = imgs + (size_t)k * N * 3;
        double* gi = gimgs + (size_t)k * N * 3;
        for (long i = 0; i < N; ++i)
            for (int c = 0; c < 3; ++c) gi[3 * i + c] = gG[(long)c * N + i] * dT(im[3 * i + c]);
        // Mask pyramid: gM^_l = gM_l + reduce^T(gM^_l+1), coarsest first.
        for (int l = L - 2; l >= 0; --l) {
            const Dims& a = F.d1[l];
            const Dims& b = F.d1[l + 1];
            reduce_T(gM.data() + b.off, b.w, b.h, gM.data() + a.off, a.w, a.h);
        }
        for (long i = 0; i < N; ++i) gmh[(size_t)k * N + i] = gM[i];
    }
    // Normalisation adjoint: gm_k = (gm^_k - sum_j gm^_j m^_j) / S.
    std::vector<double> gm((size_t)K * N);
    for (long i = 0; i < N; ++i) {
        double dot = 0.0;
        for (int k = 0; k < K; ++k) dot += gmh[(size_t)k * N + i] * F.M[k][i];
        for (int k = 0; k < K; ++k) gm[(size_t)k * N + i] = (gmh[(size_t)k * N + i] - dot) / F.S[i];
    }
    // Weight adjoint: parameters and image gradients.
    double ge = 0.0, gs = 0.0, gc = 0.0, gmu = 0.0, gsig = 0.0;
    const int r = n / 2;
    const double Nw = (double)n * n;
    for (int k = 0; k < K; ++k) {
        const Terms& t = F.terms[k];
        const double* im = imgs + (size_t)k * N * 3;
        double* gi = gimgs + (size_t)k * N * 3;
        const double* g = gm.data() + (size_t)k * N;
        std::vector<double> ggray(N, 0.0);
        for (long i = 0; i < N; ++i) {
            ge += g[i] * t.E[i];
            gc += g[i] * t.sd[i];
            gs += g[i] * t.s[i];
            const double z = (t.gray[i] - p.mu) / p.sigma;
            gmu += g[i] * p.we * t.E[i] * z / p.sigma;
            gsig += g[i] * p.we * t.E[i] * z * z / p.sigma;
            ggray[i] += g[i] * p.we * t.E[i] * (-z / p.sigma);
            const double* q = im + 3 * i;
            const double mx = q[t.imx[i]], mn = q[t.imn[i]];
            const double den = mx + mn + kEpsSat;
            gi[3 * i + t.imx[i]] += g[i] * p.ws * (2.0 * mn + kEpsSat) / (den * den);
            gi[3 * i + t.imn[i]] += g[i] * p.ws * (-(2.0 * mx + kEpsSat)) / (den * den);
        }
        for (int y = r; y < h - r; ++y)
            for (int x = r; x < w - r; ++x) {
                const long i = (long)y * w + x;
                const double a = g[i] * p.wc / ((Nw - 1.0) * t.sd[i]);
                for (int yy = y - r; yy <= y + r; ++yy)
                    for (int xx = x - r; xx <= x + r; ++xx) {
                        const long q = (long)yy * w + xx;
                        ggray[q] += a * (t.gray[q] - t.mean[i]);
                    }
            }
        for (long i = 0; i < N; ++i)
            for (int c = 0; c < 3; ++c) gi[3 * i + c] += ggray[i] / 3.0;
    }
    gprm[0] = ge; gprm[1] = gs; gprm[2] = gc; gprm[3] = gmu; gprm[4] = gsig;
    return 0;
}

}  // extern "C"
