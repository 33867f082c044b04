// Relaxed differentiable enfuse for gfx950: forward terms and hand-written
// adjoint kernels, all in double.
//
// The relaxed forward (the reference's arithmetic has a fixed-point mask
// pyramid with zero gradient almost everywhere and a constant float exposure
// weight, src/enfuse.h:637, so it cannot be differentiated as it stands):
//   gray = (r + g + b) / 3, E = exp(-((gray - mu) / sigma)^2 / 2),
//   s = (max - min) / (max + min + 1e-6),
//   sd = sqrt(sum (q - mean)^2 / (N - 1) + 1e-12) over the n x n window (interior),
//   m = w_e E + w_c sd + w_s s + 1e-12,  m^ = m / sum_k m,
//   G_k = Gaussian pyramid of T(I_k), T(v) = v >= 0 ? 1 + log1p(v) : 1 / (1 - v)
//         (src/fixmath.h:165-186), M_k = Gaussian pyramid of m^_k (reduce with
//         an all-set alpha, src/pyramid.h:166-573),
//   R_l = sum_k (G_kl - expand(G_k,l+1)) M_kl, collapse (src/pyramid.h:1757-1784),
//   O = T^-1(C_0) (src/fixmath.h:320-329).
// The adjoints are the transposes of these steps.  expand^T gathers, for a
// coarse sample, the 5 x 5 fine neighbourhood at stride 2 divided by each fine
// sample's weight sum, with the expand taps (columns 1 4 6 4 1, rows
// 1 1 6 1 1: odd rows weigh their two coarse rows equally).  reduce^T gathers,
// for a fine sample, the polyphase coarse neighbours ([1 6 1] or [4 4]) of the
// coarse gradient divided by each coarse sample's in-range weight sum.
// Laplacian levels are recomputed from the Gaussian levels wherever they are
// needed, forward and backward, instead of being stored.
#include <math.h>

#include "enbl_device.h"
#include "enbl_diff.h"
#include "enbl_internal.h"
#include "pyr_common.h"

namespace enbl {
namespace {

constexpr double kEpsSat = 1e-6;
constexpr double kEpsSd = 1e-12;
constexpr double kEpsW = 1e-12;
constexpr int kTile = 256;

inline dim3 grid2(int w, int h) { return dim3((w + kTile - 1) / kTile, h); }

__device__ __forceinline__ double T(double v) { return float_to_log_pyramid(v); }
__device__ __forceinline__ double dT(double v) { return v >= 0.0 ? 1.0 / (1.0 + v) : 1.0 / ((1.0 - v) * (1.0 - v)); }
__device__ __forceinline__ double Tinv(double c) { return log_pyramid_to_float(c); }
__device__ __forceinline__ double dTinv(double c) { return c >= 1.0 ? exp(c - 1.0) : 1.0 / (c * c); }

__device__ __forceinline__ int bin5(int d) { return d == 0 ? 6 : ((d == 1 || d == -1) ? 4 : 1); }

// Horizontal / vertical sum of the in-range reduce weights of coarse index j
// over a source of n samples.
__device__ __forceinline__ int reduce_wsum(int j, int n)
{
    int s = 0;
#pragma unroll
    for (int d = -2; d <= 2; ++d) {
        const int x = 2 * j + d;
        if (x >= 0 && x < n) s += bin5(d);
    }
    return s;
}

// expand() of one fine sample (X, Y) from a coarse plane (cw x ch), no
// wraparound, in the operation order of the expand kernels.
__device__ __forceinline__ double expand_at(const double* c, long pitch, int cw, int chh, int X, int Y)
{
    const int j = X >> 1, i = Y >> 1;
    const bool odd_x = (X & 1) != 0, odd_y = (Y & 1) != 0;
    double h3[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int y = i - 1 + k;
        h3[k] = 0.0;
        if (y < 0 || y >= chh) continue;
        const double* row = c + (long)y * pitch;
        const double x = row[j];
        const double xl = j > 0 ? row[j - 1] : 0.0;
        const double xr = j + 1 < cw ? row[j + 1] : 0.0;
        double hA, hB;
        if (cw == 1) { hA = 0.0 + x * 6.0; hB = x * 4.0; }
        else if (j == 0) { hA = (0.0 + x * 6.0) + xr; hB = (x + xr) * 4.0; }
        else if (j == cw - 1) { hA = xl + x * 6.0; hB = x * 4.0; }
        else { hA = (xl + x * 6.0) + xr; hB = (x + xr) * 4.0; }
        h3[k] = odd_x ? hB : hA;
    }
    const bool more = i + 1 < chh;
    double o;
    if (!odd_y) o = more ? (h3[0] + h3[1] * 6.0) + h3[2] : h3[0] + h3[1] * 6.0;
    else o = more ? h3[1] + h3[2] : h3[1];
    return o / (double)(expand_hsum(X, cw, false, true) * expand_vsum(Y, chh));
}

// ---- forward -----------------------------------------------------------------
__global__ __launch_bounds__(kTile) void dk_terms(const double* __restrict__ img, int w, double mu, double sigma,
                                                  double* __restrict__ gray, double* __restrict__ E,
                                                  double* __restrict__ s, long tp)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const double* q = img + ((long)y * w + x) * 3;
    const double g = (q[0] + q[1] + q[2]) / 3.0;
    const double z = (g - mu) / sigma;
    int a = 0, b = 0;
#pragma unroll
    for (int c = 1; c < 3; ++c) {
        if (q[c] > q[a]) a = c;
        if (q[c] < q[b]) b = c;
    }
    const long o = (long)y * tp + x;
    gray[o] = g;
    E[o] = exp(-0.5 * z * z);
    s[o] = (q[a] - q[b]) / (q[a] + q[b] + kEpsSat);
}

__global__ __launch_bounds__(kTile) void dk_sd(const double* __restrict__ gray, int w, int h, long tp, int n,
                                               double* __restrict__ sd, double* __restrict__ mean)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const int r = n / 2;
    double v = 0.0, mu = 0.0;
    if (x >= r && x < w - r && y >= r && y < h - r) {
        const double Nw = (double)n * n;
        double sum = 0.0;
        for (int yy = y - r; yy <= y + r; ++yy)
            for (int xx = x - r; xx <= x + r; ++xx) sum += gray[(long)yy * tp + xx];
        mu = sum / Nw;
        double ss = 0.0;
        for (int yy = y - r; yy <= y + r; ++yy)
            for (int xx = x - r; xx <= x + r; ++xx) {
                const double d = gray[(long)yy * tp + xx] - mu;
                ss += d * d;
            }
        v = sqrt(ss / (Nw - 1.0) + kEpsSd);
    }
    sd[(long)y * tp + x] = v;
    mean[(long)y * tp + x] = mu;
}

__global__ __launch_bounds__(kTile) void dk_weights(DiffTerms t, double we, double ws, double wc, DiffLevel out,
                                                    double* __restrict__ S)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= t.w) return;
    const long o = (long)y * t.tp + x;
    auto weight = [&](int k) {
        double v = we * t.E[k][o];
        v = v + wc * t.sd[k][o];
        v = v + ws * t.s[k][o];
        return v + kEpsW;
    };
    double sum = 0.0;
    for (int k = 0; k < t.K; ++k) sum = sum + weight(k);
    S[o] = sum;
    for (int k = 0; k < t.K; ++k) const_cast<double*>(out.m[k])[(long)y * out.mpitch + x] = weight(k) / sum;
}

__global__ __launch_bounds__(kTile) void dk_log(const double* __restrict__ img, int w, double* __restrict__ g, long pitch,
                                                long plane)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const double* q = img + ((long)y * w + x) * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) g[c * plane + (long)y * pitch + x] = T(q[c]);
}

__global__ __launch_bounds__(kTile) void dk_merge(DiffLevel a, double* __restrict__ r, long rpitch, long rplane)
{
    const int X = blockIdx.x * kTile + threadIdx.x;
    const int Y = blockIdx.y;
    if (X >= a.fw) return;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < a.K; ++k) {
        const double mk = a.m[k][(long)Y * a.mpitch + X];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double lap = a.g[k][c * a.gplane + (long)Y * a.gpitch + X];
            if (a.cw) lap = lap - expand_at(a.gc[k] + c * a.cplane, a.cpitch, a.cw, a.ch, X, Y);
            acc[c] = acc[c] + lap * mk;
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) r[c * rplane + (long)Y * rpitch + X] = acc[c];
}

__global__ __launch_bounds__(kTile) void dk_out(const double* __restrict__ c, long pitch, long plane, int w,
                                                double* __restrict__ out)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) out[((long)y * w + x) * 3 + ch] = Tinv(c[ch * plane + (long)y * pitch + x]);
}

// ---- adjoint -----------------------------------------------------------------
__global__ __launch_bounds__(kTile) void dk_gout(const double* __restrict__ gout, const double* __restrict__ c, long pitch,
                                                 long plane, int w, double* __restrict__ gc)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        const long o = ch * plane + (long)y * pitch + x;
        gc[o] = gout[((long)y * w + x) * 3 + ch] * dTinv(c[o]);
    }
}

// expand^T at coarse (j, i): sum over the fine neighbourhood of g / D.
// `val(X, Y)` returns the fine gradient.
template <class F>
__device__ __forceinline__ double expand_T_at(F&& val, int fw, int fh, int cw, int chh, int j, int i)
{
    double acc = 0.0;
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy) {
        const int Y = 2 * i + dy;
        if (Y < 0 || Y >= fh) continue;
        const int wy = dy == 0 ? 6 : 1;
        const int vs = expand_vsum(Y, chh);
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            const int X = 2 * j + dx;
            if (X < 0 || X >= fw) continue;
            acc += (double)(wy * bin5(dx)) * (val(X, Y) / (double)(expand_hsum(X, cw, false, true) * vs));
        }
    }
    return acc;
}

// reduce^T at fine (X, Y) of a sw x sh level from the coarse gradient plane
// (dw x dh): polyphase neighbours divided by their in-range weight sums.
__device__ __forceinline__ double reduce_T_at(const double* g, long pitch, int dw, int dh, int sw, int sh, int X, int Y)
{
    int jx[3], wx[3], nx = 0, iy[3], wy[3], ny = 0;
    const int j = X >> 1, i = Y >> 1;
    if ((X & 1) == 0) {
        if (j - 1 >= 0) { jx[nx] = j - 1; wx[nx++] = 1; }
        if (j < dw) { jx[nx] = j; wx[nx++] = 6; }
        if (j + 1 < dw) { jx[nx] = j + 1; wx[nx++] = 1; }
    } else {
        if (j < dw) { jx[nx] = j; wx[nx++] = 4; }
        if (j + 1 < dw) { jx[nx] = j + 1; wx[nx++] = 4; }
    }
    if ((Y & 1) == 0) {
        if (i - 1 >= 0) { iy[ny] = i - 1; wy[ny++] = 1; }
        if (i < dh) { iy[ny] = i; wy[ny++] = 6; }
        if (i + 1 < dh) { iy[ny] = i + 1; wy[ny++] = 1; }
    } else {
        if (i < dh) { iy[ny] = i; wy[ny++] = 4; }
        if (i + 1 < dh) { iy[ny] = i + 1; wy[ny++] = 4; }
    }
    double acc = 0.0;
    for (int a = 0; a < ny; ++a) {
        const int ws = reduce_wsum(iy[a], sh);
        for (int b = 0; b < nx; ++b)
            acc += (double)(wy[a] * wx[b]) * (g[(long)iy[a] * pitch + jx[b]] / (double)(reduce_wsum(jx[b], sw) * ws));
    }
    return acc;
}

__global__ __launch_bounds__(kTile) void dk_expand_T(const double* __restrict__ g, long gpitch, long gplane, int fw, int fh,
                                                     int C, double* __restrict__ gc, long cpitch, long cplane, int cw,
                                                     int chh)
{
    const int j = blockIdx.x * kTile + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= cw) return;
    for (int c = 0; c < C; ++c) {
        const double* gp = g + c * gplane;
        gc[c * cplane + (long)i * cpitch + j] += expand_T_at([&](int X, int Y) { return gp[(long)Y * gpitch + X]; },
                                                             fw, fh, cw, chh, j, i);
    }
}

__global__ __launch_bounds__(kTile) void dk_reduce_T(const double* __restrict__ g, long gpitch, long gplane, int dw, int dh,
                                                     int C, double* __restrict__ gf, long fpitch, long fplane, int sw,
                                                     int sh)
{
    const int X = blockIdx.x * kTile + threadIdx.x;
    const int Y = blockIdx.y;
    if (X >= sw) return;
    for (int c = 0; c < C; ++c)
        gf[c * fplane + (long)Y * fpitch + X] += reduce_T_at(g + c * gplane, gpitch, dw, dh, sw, sh, X, Y);
}

__global__ __launch_bounds__(kTile) void dk_adj_level(DiffAdjLevel a)
{
    const int X = blockIdx.x * kTile + threadIdx.x;
    const int Y = blockIdx.y;
    if (X >= a.fw) return;
    const bool coarsest = a.l == a.L - 1;
    const double m = a.M[(long)Y * a.mpitch + X];
    double gm = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double gc = a.gC[c * a.gcpl + (long)Y * a.gcp + X];
        double lap = a.G[c * a.gplane + (long)Y * a.gpitch + X];
        if (!coarsest) lap = lap - expand_at(a.Gc + c * a.gcplane, a.gcpitch, a.cw, a.ch, X, Y);
        gm += gc * lap;
        double gg = gc * m;
        if (a.l >= 1) {
            const double* gp = a.gCp + c * a.gcppl;
            gg -= expand_T_at([&](int x, int y) { return gp[(long)y * a.gcpp + x] * a.Mp[(long)y * a.mppitch + x]; },
                              a.pw, a.ph, a.fw, a.fh, X, Y);
        }
        if (!coarsest) gg += reduce_T_at(a.gGc + c * a.ggcpl, a.ggcp, a.cw, a.ch, a.fw, a.fh, X, Y);
        if (a.l == 0) {
            const long o = ((long)Y * a.fw + X) * 3 + c;
            a.gimg[o] = gg * dT(a.img[o]);
        } else {
            a.gG[c * a.ggpl + (long)Y * a.ggp + X] = gg;
        }
    }
    if (!coarsest) gm += reduce_T_at(a.gMc, a.gmcp, a.cw, a.ch, a.fw, a.fh, X, Y);
    a.gM[(long)Y * a.gmp + X] = gm;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ __launch_bounds__(kTile) void dk_weight_adj(DiffWeightAdj a)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    const bool in = x < a.w;
    double pe = 0.0, ps = 0.0, pc = 0.0, pmu = 0.0, psig = 0.0;
    if (in) {
        const long o = (long)y * a.tp + x;
        double dot = 0.0;
        for (int k = 0; k < a.K; ++k) dot += a.gmhat[k][(long)y * a.gmpitch + x] * a.mhat[k][(long)y * a.mpitch + x];
        const double S = a.S[o];
        const int r = a.n / 2;
        const bool interior = x >= r && x < a.w - r && y >= r && y < a.h - r;
        const double Nw = (double)a.n * a.n;
        for (int k = 0; k < a.K; ++k) {
            const double g = (a.gmhat[k][(long)y * a.gmpitch + x] - dot) / S;
            const double E = a.E[k][o], sd = a.sd[k][o];
            const double z = (a.gray[k][o] - a.mu) / a.sigma;
            pe += g * E;
            pc += g * sd;
            ps += g * a.s[k][o];
            pmu += g * a.we * E * z / a.sigma;
            psig += g * a.we * E * z * z / a.sigma;
            const double* q = a.img[k] + ((long)y * a.w + x) * 3;
            double* gi = a.gimg[k] + ((long)y * a.w + x) * 3;
            int ia = 0, ib = 0;
#pragma unroll
            for (int c = 1; c < 3; ++c) {
                if (q[c] > q[ia]) ia = c;
                if (q[c] < q[ib]) ib = c;
            }
            const double ge = g * a.we * E * (-z / a.sigma) / 3.0;
            const double mx = q[ia], mn = q[ib];
            const double den = mx + mn + kEpsSat;
            double gq[3] = {ge, ge, ge};
            gq[ia] += g * a.ws * (2.0 * mn + kEpsSat) / (den * den);
            gq[ib] += g * a.ws * (-(2.0 * mx + kEpsSat)) / (den * den);
#pragma unroll
            for (int c = 0; c < 3; ++c) gi[c] += gq[c];
            const double A = interior ? g * a.wc / ((Nw - 1.0) * sd) : 0.0;
            a.A[k][o] = A;
        }
    }
    pe = wave_sum(pe);
    ps = wave_sum(ps);
    pc = wave_sum(pc);
    pmu = wave_sum(pmu);
    psig = wave_sum(psig);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(a.gparam + 0, pe);
        atomicAdd(a.gparam + 1, ps);
        atomicAdd(a.gparam + 2, pc);
        atomicAdd(a.gparam + 3, pmu);
        atomicAdd(a.gparam + 4, psig);
    }
}

// Contrast adjoint of one image: d/d gray_q of sum_p A_p sd_p, i.e.
// sum over the windows p containing q of A_p (gray_q - mean_p); added to all
// three channels / 3.
__global__ __launch_bounds__(kTile) void dk_contrast_adj(const double* __restrict__ A, const double* __restrict__ mean,
                                                         const double* __restrict__ gray, int w, int h, long tp, int n,
                                                         double* __restrict__ gimg)
{
    const int x = blockIdx.x * kTile + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const int r = n / 2;
    const double gq = gray[(long)y * tp + x];
    double acc = 0.0;
    for (int yy = max(y - r, r); yy <= min(y + r, h - r - 1); ++yy)
        for (int xx = max(x - r, r); xx <= min(x + r, w - r - 1); ++xx) {
            const long p = (long)yy * tp + xx;
            acc += A[p] * (gq - mean[p]);
        }
    double* gi = gimg + ((long)y * w + x) * 3;
    const double d = acc / 3.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) gi[c] += d;
}

}  // namespace

int launch_diff_terms(const double* img, int w, int h, double mu, double sigma, double* gray, double* E, double* s,
                      long tp, hipStream_t st)
{
    hipLaunchKernelGGL(dk_terms, grid2(w, h), dim3(kTile), 0, st, img, w, mu, sigma, gray, E, s, tp);
    return 0;
}

int launch_diff_sd(const double* gray, int w, int h, long tp, int n, double* sd, double* mean, hipStream_t st)
{
    hipLaunchKernelGGL(dk_sd, grid2(w, h), dim3(kTile), 0, st, gray, w, h, tp, n, sd, mean);
    return 0;
}

int launch_diff_weights(const DiffTerms& t, double we, double ws, double wc, double* const* mhat, long mpitch,
                        double* S, hipStream_t st)
{
    DiffLevel out{};
    out.K = t.K;
    out.mpitch = mpitch;
    for (int k = 0; k < t.K; ++k) out.m[k] = mhat[k];
    hipLaunchKernelGGL(dk_weights, grid2(t.w, t.h), dim3(kTile), 0, st, t, we, ws, wc, out, S);
    return 0;
}

int launch_diff_log(const double* img, int w, int h, double* g, long pitch, long plane, hipStream_t st)
{
    hipLaunchKernelGGL(dk_log, grid2(w, h), dim3(kTile), 0, st, img, w, g, pitch, plane);
    return 0;
}

int launch_diff_merge(const DiffLevel& a, double* r, long rpitch, long rplane, hipStream_t st)
{
    hipLaunchKernelGGL(dk_merge, grid2(a.fw, a.fh), dim3(kTile), 0, st, a, r, rpitch, rplane);
    return 0;
}

int launch_diff_out(const double* c, long pitch, long plane, int w, int h, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(dk_out, grid2(w, h), dim3(kTile), 0, st, c, pitch, plane, w, out);
    return 0;
}

int launch_diff_gout(const double* gout, const double* c, long pitch, long plane, int w, int h, double* gc,
                     hipStream_t st)
{
    hipLaunchKernelGGL(dk_gout, grid2(w, h), dim3(kTile), 0, st, gout, c, pitch, plane, w, gc);
    return 0;
}

int launch_diff_expand_T(const double* g, long gpitch, long gplane, int fw, int fh, int C, double* gc, long cpitch,
                         long cplane, int cw, int ch, hipStream_t st)
{
    hipLaunchKernelGGL(dk_expand_T, grid2(cw, ch), dim3(kTile), 0, st, g, gpitch, gplane, fw, fh, C, gc, cpitch, cplane,
                       cw, ch);
    return 0;
}

int launch_diff_reduce_T(const double* g, long gpitch, long gplane, int dw, int dh, int C, double* gf, long fpitch,
                         long fplane, int sw, int sh, hipStream_t st)
{
    hipLaunchKernelGGL(dk_reduce_T, grid2(sw, sh), dim3(kTile), 0, st, g, gpitch, gplane, dw, dh, C, gf, fpitch, fplane,
                       sw, sh);
    return 0;
}

int launch_diff_adj_level(const DiffAdjLevel& a, hipStream_t st)
{
    hipLaunchKernelGGL(dk_adj_level, grid2(a.fw, a.fh), dim3(kTile), 0, st, a);
    return 0;
}

int launch_diff_weight_adj(const DiffWeightAdj& a, hipStream_t st)
{
    hipLaunchKernelGGL(dk_weight_adj, grid2(a.w, a.h), dim3(kTile), 0, st, a);
    return 0;
}

int launch_diff_contrast_adj(const double* A, const double* mean, const double* gray, int w, int h, long tp, int n,
                             double* gimg, hipStream_t st)
{
    hipLaunchKernelGGL(dk_contrast_adj, grid2(w, h), dim3(kTile), 0, st, A, mean, gray, w, h, tp, n, gimg);
    return 0;
}

}  // namespace enbl
