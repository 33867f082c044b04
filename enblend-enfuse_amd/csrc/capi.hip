// C ABI of the enblend/enfuse pyramid path (declared in include/enblend_amd.h).
//
// The reference's pyramid functions are header templates that allocate their
// levels with new and throw vigra::PreconditionViolation; here the levels live
// in one device arena per pyramid, every call is ordered on the context's
// stream and errors come back as status codes with a thread-local message.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>

#include "enbl_fused.h"
#include "enbl_tail.h"
#include "enbl_internal.h"

using namespace enbl;
using namespace enbl::host;

namespace enbl {
namespace host {

thread_local std::string g_error;
// Kernel selection (enbl_set_kernel_policy).
int g_policy = ENBL_KERNELS_AUTO;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

int hip_check(hipError_t e, const char* what)
{
    if (e == hipSuccess) return ENBL_OK;
    if (e == hipErrorOutOfMemory) return fail(ENBL_E_OOM, "%s: out of memory", what);
    return fail(ENBL_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

int launch_status(const char* what)
{
    return hip_check(hipGetLastError(), what);
}

Timed::Timed(enbl_context* c, const char* name, double bytes, double own) : ctx(c), idx((size_t)-1) {
        if (!ctx->profile) return;
        enbl_context::Timed t;
        t.name = name;
        t.bytes = bytes;
        t.own = own < 0.0 ? bytes : own;
        if (hipEventCreate(&t.start) != hipSuccess) return;
        if (hipEventCreate(&t.stop) != hipSuccess) { (void)hipEventDestroy(t.start); return; }
        (void)hipEventRecord(t.start, ctx->stream);
        ctx->timed.push_back(t);
        idx = ctx->timed.size() - 1;
}

Timed::~Timed() {
    if (idx != (size_t)-1) (void)hipEventRecord(ctx->timed[idx].stop, ctx->stream);
}

const char* pname(int t)
{
    switch (t) {
    case ENBL_I16: return "i16";
    case ENBL_I32: return "i32";
    case ENBL_F64: return "f64";
    }
    return "?";
}

size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

bool aligned2(const void* p, int ptype) { return ((uintptr_t)p % (2 * (size_t)ptype_size(ptype))) == 0; }

int mask_ptype_for(int itype) { return itype == ENBL_U8 ? ENBL_I16 : ENBL_I32; }
int image_ptype_for(int itype) { return itype == ENBL_U8 ? ENBL_I16 : itype == ENBL_U16 ? ENBL_I32 : ENBL_F64; }

int sync_streams(enbl_context* ctx)
{
    int rc = hip_check(hipStreamSynchronize(ctx->forked_from ? ctx->forked_from : ctx->stream), "stream synchronize");
    if (!rc) rc = hip_check(hipStreamSynchronize(ctx->aux_stream), "stream synchronize");
    return rc;
}

int fork_aux(enbl_context* ctx)
{
    if (ctx->forked_from) return fail(ENBL_E_ARG, "auxiliary stream already in use");
    // While the context profiles (or ENBL_SERIAL_STREAMS is set in the
    // environment, for an external profiler), the work stays on the main
    // stream: the launches then run one at a time and their durations are
    // their own, not stretched by kernels running beside them.
    static const bool serial = getenv("ENBL_SERIAL_STREAMS") != nullptr;
    if (ctx->profile || serial) return ENBL_OK;
    int rc = hip_check(hipEventRecord(ctx->aux_fork, ctx->stream), "hipEventRecord");
    if (!rc) rc = hip_check(hipStreamWaitEvent(ctx->aux_stream, ctx->aux_fork, 0), "hipStreamWaitEvent");
    if (rc) return rc;
    ctx->forked_from = ctx->stream;
    ctx->stream = ctx->aux_stream;
    return ENBL_OK;
}

int end_aux(enbl_context* ctx)
{
    if (!ctx->forked_from) return ENBL_OK;
    ctx->stream = ctx->forked_from;
    ctx->forked_from = nullptr;
    return hip_check(hipEventRecord(ctx->aux_join, ctx->aux_stream), "hipEventRecord");
}

int join_aux(enbl_context* ctx)
{
    return hip_check(hipStreamWaitEvent(ctx->stream, ctx->aux_join, 0), "hipStreamWaitEvent");
}

int alpha_scratch(enbl_context* ctx, size_t bytes, uint8_t** a, uint8_t** b)
{
    if (ctx->scratch_alpha_bytes < 2 * bytes) {
        if (ctx->scratch_alpha) {
            int rc = sync_streams(ctx);
            if (rc) return rc;
            (void)hipFree(ctx->scratch_alpha);
        }
        ctx->scratch_alpha = nullptr;
        ctx->scratch_alpha_bytes = 0;
        int rc = hip_check(hipMalloc(&ctx->scratch_alpha, 2 * bytes), "alpha scratch");
        if (rc) return rc;
        ctx->scratch_alpha_bytes = 2 * bytes;
    }
    *a = (uint8_t*)ctx->scratch_alpha;
    *b = (uint8_t*)ctx->scratch_alpha + bytes;
    return ENBL_OK;
}

void* level_ptr(const enbl_pyramid* p, int l) { return (char*)p->base + p->lv[l].offset; }

int do_reduce(enbl_context* ctx, int ptype, int C, int wrap,
              const void* src, int sw, int sh, long sp, long spl, const uint8_t* sa, long sap,
              void* dst, int dw, int dh, long dp, long dpl, uint8_t* da, long dap)
{
    if (sw < 2 || sh < 2) return fail(ENBL_E_PRECONDITION, "src image too small in reduce");
    if (dw != (sw + 1) / 2 || dh != (sh + 1) / 2)
        return fail(ENBL_E_SIZE, "reduce: destination %dx%d does not match source %dx%d", dw, dh, sw, sh);
    ReduceArgs a;
    memset(&a, 0, sizeof a);
    a.src = src; a.spitch = sp; a.splane = spl; a.sw = sw; a.sh = sh;
    a.salpha = sa; a.sapitch = sap;
    a.dst = dst; a.dpitch = dp; a.dplane = dpl; a.dw = dw; a.dh = dh;
    a.dalpha = da; a.dapitch = dap;
    a.channels = C; a.wrap = wrap; a.rows = 0; a.force_small = g_policy;
    char name[64];
    snprintf(name, sizeof name, "reduce_%s%s_c%d", sa ? "alpha_" : "", pname(ptype), C);
    const double es = (double)ptype_size(ptype) * C + (sa ? 1.0 : 0.0);
    Timed t(ctx, name, es * ((double)sw * sh + (double)dw * dh));
    int rc = launch_reduce(ptype, a, ctx->stream);
    if (rc) return fail(rc, "reduce: unsupported type/channels");
    return launch_status(name);
}

int do_expand(enbl_context* ctx, int ptype, int C, int add, int wrap,
              const void* src, int sw, int sh, long sp, long spl,
              void* dst, int dw, int dh, long dp, long dpl)
{
    if (sw < 1 || sh < 1) return fail(ENBL_E_SIZE, "expand: empty source");
    if ((dw != 2 * sw && dw != 2 * sw - 1) || (dh != 2 * sh && dh != 2 * sh - 1) || (dw < 2 && dh < 2))
        return fail(ENBL_E_SIZE, "expand: destination %dx%d does not match source %dx%d", dw, dh, sw, sh);
    ExpandArgs a;
    a.src = src; a.spitch = sp; a.splane = spl; a.sw = sw; a.sh = sh;
    a.dst = dst; a.dpitch = dp; a.dplane = dpl; a.dw = dw; a.dh = dh;
    a.channels = C; a.wrap = wrap; a.add = add; a.rows = 0; a.force_small = g_policy;
    char name[64];
    snprintf(name, sizeof name, "expand_%s_%s_c%d", add ? "add" : "sub", pname(ptype), C);
    const double es = (double)ptype_size(ptype) * C;
    Timed t(ctx, name, es * ((double)sw * sh + 2.0 * dw * dh));
    int rc = launch_expand(ptype, a, ctx->stream);
    if (rc) return fail(rc, "expand: unsupported type/channels");
    return launch_status(name);
}

int gaussian_levels(enbl_context* ctx, enbl_pyramid* p, int wrap, const uint8_t* alpha, long apitch)
{
    uint8_t* abuf[2] = {nullptr, nullptr};
    long apl = 0;
    if (alpha && p->levels > 1) {
        apl = (long)round_up((size_t)p->lv[1].w, 256);
        int rc = alpha_scratch(ctx, (size_t)apl * p->lv[1].h, &abuf[0], &abuf[1]);
        if (rc) return rc;
    }
    return gaussian_levels_with(ctx, p, wrap, alpha, apitch, abuf[0], abuf[1], apl);
}

int gaussian_levels_with(enbl_context* ctx, enbl_pyramid* p, int wrap, const uint8_t* alpha, long apitch,
                         uint8_t* a0, uint8_t* a1, long apl)
{
    uint8_t* abuf[2] = {a0, a1};
    const uint8_t* sa = alpha;
    long sap = apitch;
    for (int l = 1; l < p->levels; ++l) {
        const Level& s = p->lv[l - 1];
        const Level& d = p->lv[l];
        uint8_t* da = alpha ? abuf[(l - 1) & 1] : nullptr;
        int rc = do_reduce(ctx, p->ptype, p->channels, wrap, level_ptr(p, l - 1), s.w, s.h, s.pitch, s.plane,
                           sa, sap, level_ptr(p, l), d.w, d.h, d.pitch, d.plane, da, apl);
        if (rc) return rc;
        sa = da;
        sap = apl;
    }
    return ENBL_OK;
}

int laplacian_levels(enbl_context* ctx, enbl_pyramid* p, int wrap)
{
    for (int l = 0; l + 1 < p->levels; ++l) {
        const Level& f = p->lv[l];
        const Level& c = p->lv[l + 1];
        int rc = do_expand(ctx, p->ptype, p->channels, 0, wrap, level_ptr(p, l + 1), c.w, c.h, c.pitch, c.plane,
                           level_ptr(p, l), f.w, f.h, f.pitch, f.plane);
        if (rc) return rc;
    }
    return ENBL_OK;
}

int build_level0(enbl_context* ctx, enbl_pyramid* p, int itype, int frac, const void* img, long ipitch)
{
    const Level& l0 = p->lv[0];
    char name[64];
    snprintf(name, sizeof name, "to_pyramid_%s_c%d", pname(p->ptype), p->channels);
    Timed t(ctx, name, (double)l0.w * l0.h * p->channels * (ptype_size(itype) + ptype_size(p->ptype)));
    int rc = launch_to_pyramid(itype, p->ptype, frac, p->channels, img, ipitch, l0.w, l0.h,
                               level_ptr(p, 0), l0.pitch, l0.plane, ctx->stream);
    if (rc) return fail(rc, "copyToPyramidImage: unsupported image/pyramid type pair");
    return launch_status(name);
}

int cached_pyramid(enbl_context* ctx, int role, int ptype, int C, int w, int h, int levels, int slot,
                   enbl_pyramid** out)
{
    int seen = 0;
    for (auto& c : ctx->cache) {
        if (c.role == role && c.ptype == ptype && c.channels == C && c.w == w && c.h == h && c.levels == levels) {
            if (seen == slot) { *out = c.p; return ENBL_OK; }
            ++seen;
        }
    }
    enbl_pyramid* p = nullptr;
    int rc = enbl_pyramid_create(ctx, ptype, C, w, h, levels, &p);
    if (rc) return rc;
    ctx->cache.push_back({role, ptype, C, w, h, levels, p});
    *out = p;
    return ENBL_OK;
}

bool streaming_level(int cw, int ch)
{
    if (cw < 2 || ch < 2) return false;
    if (g_policy == ENBL_KERNELS_STREAMING) return true;
    if (g_policy == ENBL_KERNELS_DIRECT) return false;
    return (long)cw * ch > (1L << 20);
}

int alpha_pingpong(enbl_context* ctx, const enbl_pyramid* p, uint8_t** a, uint8_t** b, long* pitch)
{
    *pitch = (long)round_up((size_t)p->lv[1].w, 256);
    return alpha_scratch(ctx, (size_t)*pitch * p->lv[1].h, a, b);
}

// Number of leading levels built by the fused reduce + expand-subtract
// pass: those whose coarse level is a streaming level (0 with wraparound or
// without alpha).
int fused_level_count(const enbl_pyramid* p, int wrap, const uint8_t* alpha)
{
    int s = 0;
    if (!wrap && alpha)
        while (s + 1 < p->levels && streaming_level(p->lv[s + 1].w, p->lv[s + 1].h)) ++s;
    return s;
}

// Fused passes over levels l0..s-1 (s >= 1): Laplacian levels l0..s-1 and
// Gaussian level s land in the pyramid; the alpha of level l+1 goes to
// abuf[l & 1] (pitch apl), so level s's alpha ends in abuf[(s - 1) & 1].
// l0 = 1 starts from Gaussian level 1 in g1 (level-1 pitch and plane) with
// its alpha in abuf[0]; level 0 of the pyramid is then not written.
int fused_laplacian_levels(enbl_context* ctx, int itype, int frac, const void* img, long ipitch,
                           const uint8_t* alpha, long apitch, enbl_pyramid* p, int s, uint8_t* const abuf[2], long apl,
                           int l0 = 0, const void* g1 = nullptr)
{
    const size_t es = (size_t)ptype_size(p->ptype);
    void* gbuf[2] = {nullptr, nullptr};
    if (s >= 2) {
        const size_t gbytes = round_up((size_t)p->lv[1].plane * p->channels * es, 256);
        void* g;
        int rc = scratch(ctx, 8, 2 * gbytes, &g);
        if (rc) return rc;
        gbuf[0] = g;
        gbuf[1] = (char*)g + gbytes;
    }
    // Float images: the log transform of level 0 is made once per sample by a
    // pointwise pass into a double scratch plane; inside the fused pass every
    // fine row would be transformed twice (its horizontal sums and its
    // Laplacian), and the transform costs far more than the extra bytes.
    const void* g0 = nullptr;
    if (l0 == 0 && itype == ENBL_F32) {
        const Level& F = p->lv[0];
        void* g;
        int rc = scratch(ctx, 23, round_up((size_t)F.plane * p->channels * es, 256), &g);
        if (rc) return rc;
        char name[64];
        snprintf(name, sizeof name, "to_pyramid_%s_c%d", pname(p->ptype), p->channels);
        Timed t(ctx, name, (double)F.w * F.h * p->channels * (ptype_size(itype) + (double)es));
        rc = launch_to_pyramid(itype, p->ptype, frac, p->channels, img, ipitch, F.w, F.h, g, F.pitch, F.plane,
                               ctx->stream);
        if (rc) return fail(rc, "copyToPyramidImage: unsupported image/pyramid type pair");
        rc = launch_status(name);
        if (rc) return rc;
        g0 = g;
    }
    for (int l = l0; l < s; ++l) {
        const Level& F = p->lv[l];
        const Level& Cl = p->lv[l + 1];
        FusedLapArgs a;
        if (l == 0 && g0) { a.src = g0; a.spitch = F.pitch; a.splane = F.plane; a.salpha = alpha; a.sapitch = apitch; }
        else if (l == 0) { a.src = img; a.spitch = ipitch; a.splane = 0; a.salpha = alpha; a.sapitch = apitch; }
        else { a.src = l == 1 && g1 ? g1 : gbuf[(l - 1) & 1]; a.spitch = F.pitch; a.splane = F.plane; a.salpha = abuf[(l - 1) & 1]; a.sapitch = apl; }
        a.frac = frac;
        a.fw = F.w; a.fh = F.h;
        a.lap = level_ptr(p, l); a.lpitch = F.pitch; a.lplane = F.plane;
        if (l + 1 < s) { a.coarse = gbuf[l & 1]; }
        else { a.coarse = level_ptr(p, l + 1); }
        a.cpitch = Cl.pitch; a.cplane = Cl.plane;
        a.calpha = abuf[l & 1]; a.capitch = apl;
        a.cw = Cl.w; a.ch = Cl.h;
        a.rows = 0;
        const double esC = (double)es * p->channels;
        char name[64];
        snprintf(name, sizeof name, "laplacian_level%s_%s_c%d", l == 0 ? "0" : "", pname(p->ptype), p->channels);
        // credited bytes: the reduce and the expand-subtract it replaces
        // (and the level-0 conversion); own bytes: the source level and its
        // alpha read once, the Laplacian level and the coarse level with its
        // alpha written once
        const double nF = (double)F.w * F.h, nC = (double)Cl.w * Cl.h;
        const bool from_image = l == 0 && !g0;
        const double srcC = from_image ? (double)p->channels * ptype_size(itype) : esC;
        Timed t(ctx, name, (esC + 1.0) * (nF + nC) + esC * (nC + 2.0 * nF) +
                               (from_image ? nF * p->channels * (ptype_size(itype) + (double)es) : 0.0),
                nF * (srcC + 1.0) + nF * esC + nC * (esC + 1.0));
        int rc = launch_laplacian_level(p->ptype, from_image ? itype : 0, a, p->channels, ctx->stream);
        if (rc) return fail(rc, "laplacian level: unsupported type");
        rc = launch_status(name);
        if (rc) return rc;
    }
    return ENBL_OK;
}

// laplacianPyramid (src/pyramid.h:1646-1747) with the fused
// reduce + expand-subtract pass on every level whose coarse level is large,
// and the separate reduce / expand kernels on the small tail.
int laplacian_pyramid_any(enbl_context* ctx, int wrap, int itype, int frac, const void* img, long ipitch,
                          const uint8_t* alpha, long apitch, enbl_pyramid* p)
{
    const int L = p->levels;
    const int s = fused_level_count(p, wrap, alpha);
    if (s == 0) {
        int rc = build_level0(ctx, p, itype, frac, img, ipitch);
        if (!rc) rc = gaussian_levels(ctx, p, wrap, alpha, apitch);
        if (!rc) rc = laplacian_levels(ctx, p, wrap);
        return rc;
    }
    uint8_t* abuf[2];
    long apl;
    int rc = alpha_pingpong(ctx, p, &abuf[0], &abuf[1], &apl);
    if (!rc) rc = fused_laplacian_levels(ctx, itype, frac, img, ipitch, alpha, apitch, p, s, abuf, apl);
    if (rc) return rc;
    // Small tail: Gaussian levels s+1.. from level s in place, then
    // expand-subtract levels s..L-2.
    const uint8_t* sa = abuf[(s - 1) & 1];
    for (int l = s + 1; l < L; ++l) {
        const Level& S_ = p->lv[l - 1];
        const Level& D = p->lv[l];
        uint8_t* da = abuf[(l - 1) & 1];
        rc = do_reduce(ctx, p->ptype, p->channels, wrap, level_ptr(p, l - 1), S_.w, S_.h, S_.pitch, S_.plane,
                       sa, apl, level_ptr(p, l), D.w, D.h, D.pitch, D.plane, da, apl);
        if (rc) return rc;
        sa = da;
    }
    for (int l = s; l + 1 < L; ++l) {
        const Level& F = p->lv[l];
        const Level& Cl = p->lv[l + 1];
        rc = do_expand(ctx, p->ptype, p->channels, 0, wrap, level_ptr(p, l + 1), Cl.w, Cl.h, Cl.pitch, Cl.plane,
                       level_ptr(p, l), F.w, F.h, F.pitch, F.plane);
        if (rc) return rc;
    }
    return ENBL_OK;
}

// One multi-job launch of small-level work (tail.hip).
struct TailBuilder {
    TailLaunch t;
    double bytes = 0.0;
    TailBuilder() { memset(&t, 0, sizeof t); }
    TailJob& add(int kind, int w, int h) { return add_tiles(kind, (w + 63) / 64, (h + 3) / 4); }
    TailJob& add_tiles(int kind, int tx, int ty) {
        TailJob& J = t.job[t.njobs++];
        J.kind = kind;
        J.tiles_x = tx;
        J.block0 = t.blocks;
        t.blocks += tx * ty;
        return J;
    }
    // A reduce job: streaming strips for outputs of at least kTailStripPixels
    // (their 5x5 windows would load every source sample 6 times), one thread
    // per output below.
    void add_reduce(bool image, ReduceArgs a) { add_reduce_kind(image ? kTailReduceImage : kTailReduceMask, a); }
    // kind: kTailReduceImage, kTailReduceMask or kTailReduceMaskAlpha
    void add_reduce_kind(int kind, ReduceArgs a) {
        if ((long)a.dw * a.dh >= kTailStripPixels) {
            const int waves_x = (a.dw + kOwned - 1) / kOwned + 1;
            a.rows = 4;
            const int strip = kind == kTailReduceImage ? kTailReduceImageStrip
                            : (kind == kTailReduceMask ? kTailReduceMaskStrip : kTailReduceMaskAlphaStrip);
            add_tiles(strip, (waves_x + 3) / 4, (a.dh + a.rows - 1) / a.rows).u.r = a;
        } else {
            add(kind, a.dw, a.dh).u.r = a;
        }
    }
    bool full() const { return t.njobs == kTailMaxJobs; }
    static constexpr long kTailStripPixels = 1L << 16;
};

ReduceArgs tail_reduce(const enbl_pyramid* p, int l, const uint8_t* sa, long sap, uint8_t* da, long dap)
{
    ReduceArgs a;
    memset(&a, 0, sizeof a);
    const Level& S_ = p->lv[l];
    const Level& D = p->lv[l + 1];
    a.src = level_ptr(p, l); a.spitch = S_.pitch; a.splane = S_.plane; a.sw = S_.w; a.sh = S_.h;
    a.salpha = sa; a.sapitch = sap;
    a.dst = level_ptr(p, l + 1); a.dpitch = D.pitch; a.dplane = D.plane; a.dw = D.w; a.dh = D.h;
    a.dalpha = da; a.dapitch = dap;
    a.channels = p->channels;
    return a;
}

BlendArgs tail_blend(const enbl_pyramid* m, const enbl_pyramid* wlp, enbl_pyramid* blp, int l, double wv)
{
    BlendArgs b;
    memset(&b, 0, sizeof b);
    b.mask = level_ptr(m, l); b.mpitch = m->lv[l].pitch;
    b.white = level_ptr(wlp, l); b.wpitch = wlp->lv[l].pitch; b.wplane = wlp->lv[l].plane;
    b.black = level_ptr(blp, l); b.bpitch = blp->lv[l].pitch; b.bplane = blp->lv[l].plane;
    b.w = blp->lv[l].w; b.h = blp->lv[l].h;
    b.white_value = wv;
    return b;
}

int run_tail(enbl_context* ctx, int pt, int mt, int C, TailBuilder& b)
{
    char name[64];
    snprintf(name, sizeof name, "tail_%s_c%d", pname(pt), C);
    Timed t(ctx, name, b.bytes);
    int rc = launch_tail(pt, mt, C, b.t, ctx->stream);
    if (rc) return fail(rc, "small-level launch: unsupported types");
    return launch_status(name);
}

// The enblend step below the fused levels (no wraparound): the white and
// black Laplacian tails and the mask Gaussian tail from level s (Gaussian
// levels s in place, alphas of level s given), then the blend of every
// level >= s and the collapse down to level s.  Launch t reduces level s+t
// of all three pyramids and expand-subtracts level s+t-1 of the two image
// pyramids: no job of a launch writes what another reads.
int enblend_tail(enbl_context* ctx, const enbl_pyramid* mgp, enbl_pyramid* wlp, enbl_pyramid* blp, int s,
                 const uint8_t* wa_s, const uint8_t* ba_s, long apl, double wv)
{
    const int L = blp->levels;
    const int pt = blp->ptype, mt = mgp->ptype, C = blp->channels;
    const double esP = (double)ptype_size(pt) * C, esM = ptype_size(mt);
    // alpha of the tail levels s+1..L-1, white then black
    std::vector<size_t> aoff(L + 1, 0);
    size_t abytes = 0;
    for (int l = s + 1; l < L; ++l) {
        aoff[l] = abytes;
        abytes += round_up((size_t)round_up((size_t)blp->lv[l].w, 64) * blp->lv[l].h, 256);
    }
    uint8_t* tail_alpha = nullptr;
    if (abytes) {
        void* p;
        int rc = scratch(ctx, 15, 2 * abytes, &p);
        if (rc) return rc;
        tail_alpha = (uint8_t*)p;
    }
    auto apitch = [&](int l) { return l == s ? apl : (long)round_up((size_t)blp->lv[l].w, 64); };
    auto walpha = [&](int l) -> uint8_t* { return l == s ? const_cast<uint8_t*>(wa_s) : tail_alpha + aoff[l]; };
    auto balpha = [&](int l) -> uint8_t* { return l == s ? const_cast<uint8_t*>(ba_s) : tail_alpha + abytes + aoff[l]; };
    auto expand_sub = [&](TailBuilder& b, enbl_pyramid* p, int l) {
        const Level& F = p->lv[l];
        const Level& Cl = p->lv[l + 1];
        ExpandArgs e;
        memset(&e, 0, sizeof e);
        e.src = level_ptr(p, l + 1); e.spitch = Cl.pitch; e.splane = Cl.plane; e.sw = Cl.w; e.sh = Cl.h;
        e.dst = level_ptr(p, l); e.dpitch = F.pitch; e.dplane = F.plane; e.dw = F.w; e.dh = F.h;
        e.channels = C;
        b.add(kTailExpandSub, Cl.w, Cl.h).u.e = e;
        b.bytes += esP * ((double)Cl.w * Cl.h + 2.0 * F.w * F.h);
    };
    for (int l = s; l + 1 < L; ++l) {
        const Level& S_ = blp->lv[l];
        const Level& D = blp->lv[l + 1];
        if (S_.w < 2 || S_.h < 2) return fail(ENBL_E_PRECONDITION, "src image too small in reduce");
        TailBuilder b;
        b.add_reduce(true, tail_reduce(wlp, l, walpha(l), apitch(l), walpha(l + 1), apitch(l + 1)));
        b.add_reduce(true, tail_reduce(blp, l, balpha(l), apitch(l), balpha(l + 1), apitch(l + 1)));
        b.add_reduce(false, tail_reduce(mgp, l, nullptr, 0, nullptr, 0));
        b.bytes += 2.0 * (esP + 1.0) * ((double)S_.w * S_.h + (double)D.w * D.h) + esM * ((double)S_.w * S_.h + (double)D.w * D.h);
        if (l > s) {
            expand_sub(b, wlp, l - 1);
            expand_sub(b, blp, l - 1);
        }
        int rc = run_tail(ctx, pt, mt, C, b);
        if (rc) return rc;
    }
    if (L - 2 >= s) {
        TailBuilder b;
        expand_sub(b, wlp, L - 2);
        expand_sub(b, blp, L - 2);
        int rc = run_tail(ctx, pt, mt, C, b);
        if (rc) return rc;
    }
    // Blend of the coarsest level (src/enblend.h:472-474), then blend +
    // collapse (src/pyramid.h:1757-1784) of levels L-2..s.
    {
        TailBuilder b;
        const Level& T = blp->lv[L - 1];
        b.add(kTailBlend, T.w, T.h).u.b = tail_blend(mgp, wlp, blp, L - 1, wv);
        b.bytes += (double)T.w * T.h * (esM + 3.0 * esP);
        int rc = run_tail(ctx, pt, mt, C, b);
        if (rc) return rc;
    }
    for (int l = L - 2; l >= s; --l) {
        TailBuilder b;
        const Level& F = blp->lv[l];
        const Level& Cl = blp->lv[l + 1];
        CollapseSmallArgs c;
        memset(&c, 0, sizeof c);
        c.coarse = level_ptr(blp, l + 1); c.cpitch = Cl.pitch; c.cplane = Cl.plane; c.cw = Cl.w; c.ch = Cl.h;
        c.fine = tail_blend(mgp, wlp, blp, l, wv);
        b.add(kTailCollapse, Cl.w, Cl.h).u.c = c;
        b.bytes += (double)F.w * F.h * (esM + 3.0 * esP) + esP * ((double)Cl.w * Cl.h + 2.0 * F.w * F.h);
        int rc = run_tail(ctx, pt, mt, C, b);
        if (rc) return rc;
    }
    return ENBL_OK;
}

namespace {

// Multi-job launches filled job by job: a launch goes out when it holds
// kTailMaxJobs jobs and at flush().  pt / mt / C: the pyramid types and
// image channels the launch is instantiated for.
struct TailSink {
    enbl_context* ctx;
    int pt, mt, C;
    TailBuilder b;
    TailSink(enbl_context* c, int p, int m, int ch) : ctx(c), pt(p), mt(m), C(ch) {}
    int room() {
        if (!b.full()) return ENBL_OK;
        return flush();
    }
    int flush() {
        int rc = b.t.njobs ? run_tail(ctx, pt, mt, C, b) : ENBL_OK;
        b = TailBuilder();
        return rc;
    }
    int reduce(int kind, const ReduceArgs& a, double bytes) {
        int rc = room();
        if (rc) return rc;
        b.add_reduce_kind(kind, a);
        b.bytes += bytes;
        return ENBL_OK;
    }
    int expand_sub(const enbl_pyramid* p, int l) {
        int rc = room();
        if (rc) return rc;
        const Level& F = p->lv[l];
        const Level& Cl = p->lv[l + 1];
        ExpandArgs e;
        memset(&e, 0, sizeof e);
        e.src = level_ptr(p, l + 1); e.spitch = Cl.pitch; e.splane = Cl.plane; e.sw = Cl.w; e.sh = Cl.h;
        e.dst = level_ptr(p, l); e.dpitch = F.pitch; e.dplane = F.plane; e.dw = F.w; e.dh = F.h;
        e.channels = p->channels;
        b.add(kTailExpandSub, Cl.w, Cl.h).u.e = e;
        b.bytes += (double)ptype_size(p->ptype) * p->channels * ((double)Cl.w * Cl.h + 2.0 * F.w * F.h);
        return ENBL_OK;
    }
};

// The pyramid-type pair a multi-job launch is instantiated for: image and
// mask pyramids of the same image type.
int tail_mask_type(int pt) { return pt == ENBL_I16 ? ENBL_I16 : ENBL_I32; }
int tail_image_type(int mt) { return mt == ENBL_I16 ? ENBL_I16 : ENBL_I32; }

}  // namespace

// Laplacian pyramids of K images under their alphas, no wraparound.  The
// levels whose coarse level streams are built image by image (the fused
// Laplacian passes); the levels below, of all K pyramids, go through shared
// multi-job launches: the launch for level l reduces level l of every
// pyramid and expand-subtracts its level l-1, split into launches of at most
// kTailMaxJobs jobs (no job of a level writes what another job of that level
// reads).  Same results as laplacian_pyramid_any per image.  Scratch: the
// alpha ping-pong, slots 8 and 23 (fused passes) and 35 (alpha levels).
int enfuse_level0_count(const enbl_pyramid* p, const uint8_t* alpha) { return fused_level_count(p, 0, alpha); }

int enfuse_image_pyramids(enbl_context* ctx, int itype, int frac, int K, const void* const* imgs, const long* ips,
                          const uint8_t* const* alphas, const long* aps, enbl_pyramid* const* lps, void* const* g1s)
{
    const enbl_pyramid* p0 = lps[0];
    const int L = p0->levels;
    const int pt = p0->ptype, C = p0->channels;
    const int s = fused_level_count(p0, 0, alphas[0]);
    const double esP = (double)ptype_size(pt) * C;
    auto tpitch = [&](int l) { return (long)round_up((size_t)p0->lv[l].w, 64); };
    // Per image: its level-s alpha as the fused passes leave it (pitch apl,
    // sized for every level those passes write into that ping-pong half),
    // then its levels s+1..L-1.
    uint8_t* pp[2] = {nullptr, nullptr};
    long apl = 0;
    size_t own_s = 0;
    if (s >= 1) {
        int rc = alpha_pingpong(ctx, p0, &pp[0], &pp[1], &apl);
        if (rc) return rc;
        for (int l = s; l >= 1; l -= 2) own_s = std::max(own_s, (size_t)apl * p0->lv[l].h);
        own_s = round_up(own_s, 256);
    }
    std::vector<size_t> toff(L, 0);
    size_t tbytes = own_s;
    for (int l = s + 1; l < L; ++l) {
        toff[l] = tbytes;
        tbytes += round_up((size_t)tpitch(l) * p0->lv[l].h, 256);
    }
    uint8_t* base = nullptr;
    if (tbytes) {
        void* v;
        int rc = scratch(ctx, 35, tbytes * (size_t)K, &v);
        if (rc) return rc;
        base = (uint8_t*)v;
    }
    auto ialpha = [&](int m, int l) -> uint8_t* {
        if (l == 0) return const_cast<uint8_t*>(alphas[m]);
        return base + tbytes * (size_t)m + (l == s ? 0 : toff[l]);
    };
    auto ipitch = [&](int m, int l) { return l == 0 ? aps[m] : (l == s ? apl : tpitch(l)); };
    if (g1s && s < 1) return fail(ENBL_E_ARG, "enfuse_image_pyramids: level 1 is not a streaming level");
    for (int m = 0; m < K; ++m) {
        int rc;
        if (s == 0) {
            rc = build_level0(ctx, lps[m], itype, frac, imgs[m], ips[m]);
        } else if (g1s) {
            // Gaussian level 1 (and its alpha) reduced from the image, kept
            // apart for the level-0 merge, which forms Laplacian level 0 on
            // the fly; the fused passes go on from level 1.
            uint8_t* abuf[2] = {pp[0], pp[1]};
            abuf[(s - 1) & 1] = ialpha(m, s);
            const Level& F = p0->lv[0];
            const Level& D = p0->lv[1];
            ReduceArgs r;
            memset(&r, 0, sizeof r);
            r.src = imgs[m]; r.spitch = ips[m]; r.splane = 0; r.sw = F.w; r.sh = F.h;
            r.salpha = alphas[m]; r.sapitch = aps[m];
            r.dst = g1s[m]; r.dpitch = D.pitch; r.dplane = D.plane; r.dw = D.w; r.dh = D.h;
            r.dalpha = abuf[0]; r.dapitch = apl;
            r.channels = C; r.frac = frac;
            // a float image's level 0 (its log transform) is also written, into
            // the pyramid's level 0, for the level-0 merge to read
            const bool conv = itype == ENBL_F32;
            if (conv) { r.conv = level_ptr(lps[m], 0); r.cpitch = F.pitch; r.cplane = F.plane; }
            char name[64];
            snprintf(name, sizeof name, "reduce_image0_%s_c%d", pname(pt), C);
            const double n0 = (double)F.w * F.h, n1 = (double)D.w * D.h;
            Timed t(ctx, name, n0 * (C * (double)ptype_size(itype) + 1.0 + (conv ? esP : 0.0)) + n1 * (esP + 1.0));
            rc = launch_reduce_image(itype, pt, r, ctx->stream);
            if (rc) return fail(rc, "level-0 reduce: unsupported type");
            rc = launch_status(name);
            if (!rc && s >= 2)
                rc = fused_laplacian_levels(ctx, itype, frac, imgs[m], ips[m], alphas[m], aps[m], lps[m], s, abuf, apl, 1,
                                            g1s[m]);
            else if (!rc)
                rc = hip_check(hipMemcpyAsync(level_ptr(lps[m], 1), g1s[m], (size_t)D.plane * C * ptype_size(pt),
                                              hipMemcpyDeviceToDevice, ctx->stream), "level 1");
        } else {
            uint8_t* abuf[2] = {pp[0], pp[1]};
            abuf[(s - 1) & 1] = ialpha(m, s);
            rc = fused_laplacian_levels(ctx, itype, frac, imgs[m], ips[m], alphas[m], aps[m], lps[m], s, abuf, apl);
        }
        if (rc) return rc;
    }
    // Launch t reduces level s+t of every pyramid and expand-subtracts its
    // level s+t-1.
    TailSink sink(ctx, pt, tail_mask_type(pt), C);
    for (int l = s; l < L; ++l) {
        for (int m = 0; m < K; ++m) {
            enbl_pyramid* p = lps[m];
            int rc = ENBL_OK;
            if (l + 1 < L) {
                if (p->lv[l].w < 2 || p->lv[l].h < 2) return fail(ENBL_E_PRECONDITION, "src image too small in reduce");
                rc = sink.reduce(kTailReduceImage,
                                 tail_reduce(p, l, ialpha(m, l), ipitch(m, l), ialpha(m, l + 1), ipitch(m, l + 1)),
                                 (esP + 1.0) * ((double)p->lv[l].w * p->lv[l].h + (double)p->lv[l + 1].w * p->lv[l + 1].h));
            }
            if (!rc && l > s) rc = sink.expand_sub(p, l - 1);
            if (rc) return rc;
        }
        int rc = sink.flush();
        if (rc) return rc;
    }
    return ENBL_OK;
}

// Gaussian pyramids of K masks (level 0 written) under one shared alpha `ua`,
// no wraparound: the streaming levels mask by mask (the first writes the
// alpha's levels), the levels below of all K in shared multi-job launches.
// Same results as gaussian_levels per mask.  Scratch: slot 36 (the alpha's
// levels).
int enfuse_mask_pyramids(enbl_context* ctx, int K, enbl_pyramid* const* mgps, const uint8_t* ua, long uap,
                         const MaskLevel0* l0)
{
    const enbl_pyramid* g0 = mgps[0];
    const int L = g0->levels;
    const int mt = g0->ptype;
    const int s = fused_level_count(g0, 0, ua);
    const double esM = ptype_size(mt);
    auto tpitch = [&](int l) { return (long)round_up((size_t)g0->lv[l].w, 64); };
    std::vector<size_t> uoff(L, 0);
    size_t ubytes = 0;
    for (int l = 1; l < L; ++l) {
        uoff[l] = ubytes;
        ubytes += round_up((size_t)tpitch(l) * g0->lv[l].h, 256);
    }
    uint8_t* base = nullptr;
    if (ubytes) {
        void* v;
        int rc = scratch(ctx, 36, ubytes, &v);
        if (rc) return rc;
        base = (uint8_t*)v;
    }
    auto ualpha = [&](int l) -> uint8_t* { return l == 0 ? const_cast<uint8_t*>(ua) : base + uoff[l]; };
    auto upitch = [&](int l) { return l == 0 ? uap : tpitch(l); };
    if (l0 && s < 1) return fail(ENBL_E_ARG, "enfuse_mask_pyramids: level 1 is not a streaming level");
    for (int m = 0; m < K; ++m) {
        enbl_pyramid* g = mgps[m];
        for (int l = 0; l < s; ++l) {
            const Level& S_ = g->lv[l];
            const Level& D = g->lv[l + 1];
            int rc;
            if (l == 0 && l0) {
                // level 0 formed from the weights and the norm on the fly
                NormReduceArgs a;
                memset(&a, 0, sizeof a);
                a.src = l0->w[m]; a.spitch = l0->wpitch; a.splane = 0; a.sw = S_.w; a.sh = S_.h;
                a.salpha = ua; a.sapitch = uap;
                a.dst = level_ptr(g, 1); a.dpitch = D.pitch; a.dplane = D.plane; a.dw = D.w; a.dh = D.h;
                a.dalpha = m == 0 ? ualpha(1) : nullptr; a.dapitch = upitch(1);
                a.channels = 1;
                a.norm = l0->norm; a.npitch = l0->npitch;
                a.scale = l0->scale; a.fallback = l0->fallback; a.hard = l0->hard;
                char name[64];
                snprintf(name, sizeof name, "reduce_mask0_norm_%s", pname(mt));
                Timed t(ctx, name, (double)S_.w * S_.h * 9.0 + (double)D.w * D.h * (esM + (m == 0 ? 1.0 : 0.0)));
                rc = launch_reduce_norm_mask(mt, a, ctx->stream);
                if (rc) return fail(rc, "mask level-0 reduce: unsupported type");
                rc = launch_status(name);
            } else {
                rc = do_reduce(ctx, mt, 1, 0, level_ptr(g, l), S_.w, S_.h, S_.pitch, S_.plane, ualpha(l), upitch(l),
                               level_ptr(g, l + 1), D.w, D.h, D.pitch, D.plane, m == 0 ? ualpha(l + 1) : nullptr,
                               upitch(l + 1));
            }
            if (rc) return rc;
        }
    }
    TailSink sink(ctx, tail_image_type(mt), mt, 1);
    for (int l = s; l + 1 < L; ++l) {
        for (int m = 0; m < K; ++m) {
            enbl_pyramid* g = mgps[m];
            if (g->lv[l].w < 2 || g->lv[l].h < 2) return fail(ENBL_E_PRECONDITION, "src image too small in reduce");
            int rc = sink.reduce(kTailReduceMaskAlpha,
                                 tail_reduce(g, l, ualpha(l), upitch(l), m == 0 ? ualpha(l + 1) : nullptr, upitch(l + 1)),
                                 (esM + 1.0) * (double)g->lv[l].w * g->lv[l].h +
                                     (esM + (m == 0 ? 1.0 : 0.0)) * (double)g->lv[l + 1].w * g->lv[l + 1].h);
            if (rc) return rc;
        }
        int rc = sink.flush();
        if (rc) return rc;
    }
    return ENBL_OK;
}

int scratch(enbl_context* ctx, int slot, size_t bytes, void** out)
{
    if (slot < 0 || slot >= kScratchSlots) return fail(ENBL_E_ARG, "scratch slot %d", slot);
    if (ctx->scratch_bytes[slot] < bytes) {
        if (ctx->scratch_buf[slot]) {
            int rc = sync_streams(ctx);
            if (rc) return rc;
            (void)hipFree(ctx->scratch_buf[slot]);
            ctx->scratch_buf[slot] = nullptr;
            ctx->scratch_bytes[slot] = 0;
        }
        int rc = hip_check(hipMalloc(&ctx->scratch_buf[slot], bytes), "scratch allocation");
        if (rc) return rc;
        ctx->scratch_bytes[slot] = bytes;
    }
    *out = ctx->scratch_buf[slot];
    return ENBL_OK;
}

int collapse_levels(enbl_context* ctx, int wrap, enbl_pyramid* p, int lmin)
{
    for (int l = p->levels - 2; l >= lmin; --l) {
        const Level& f = p->lv[l];
        const Level& c = p->lv[l + 1];
        int rc = do_expand(ctx, p->ptype, p->channels, 1, wrap, level_ptr(p, l + 1), c.w, c.h, c.pitch, c.plane,
                           level_ptr(p, l), f.w, f.h, f.pitch, f.plane);
        if (rc) return rc;
    }
    return ENBL_OK;
}

}  // namespace host

std::vector<Level> plan_levels(int ptype, int C, int w, int h, int levels, size_t* bytes)
{
    std::vector<Level> lv;
    size_t off = 0;
    const size_t es = (size_t)ptype_size(ptype);
    for (int l = 0; l < levels; ++l) {
        Level L;
        L.w = w;
        L.h = h;
        L.pitch = (long)round_up((size_t)w, 64);
        L.plane = (long)round_up((size_t)L.pitch * h, 64);
        L.offset = off;
        off += round_up((size_t)L.plane * C * es, 256);
        lv.push_back(L);
        w = (w + 1) >> 1;
        h = (h + 1) >> 1;
    }
    if (bytes) *bytes = off;
    return lv;
}

}  // namespace enbl

extern "C" {

const char* enbl_last_error(void) { return g_error.c_str(); }

int enbl_context_create(int device, enbl_context** out)
{
    if (!out) return fail(ENBL_E_ARG, "enbl_context_create: null out");
    int n = 0;
    int rc = hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (rc) return rc;
    if (device < 0 || device >= n) return fail(ENBL_E_ARG, "enbl_context_create: no HIP device %d", device);
    rc = hip_check(hipSetDevice(device), "hipSetDevice");
    if (rc) return rc;
    enbl_context* c = new enbl_context();
    c->device = device;
    c->profile = false;
    c->scratch_alpha = nullptr;
    c->scratch_alpha_bytes = 0;
    for (int i = 0; i < kScratchSlots; ++i) { c->scratch_buf[i] = nullptr; c->scratch_bytes[i] = 0; }
    c->batch = nullptr;
    for (int i = 0; i < 3; ++i) { c->batch_rows[i] = 0; c->batch_bpc[i] = 0; }
    c->colorspace = ENBL_COLORSPACE_IDENTITY;
    rc = hip_check(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking), "hipStreamCreate");
    if (rc) { delete c; return rc; }
    rc = hip_check(hipStreamCreateWithFlags(&c->aux_stream, hipStreamNonBlocking), "hipStreamCreate");
    if (rc) { (void)hipStreamDestroy(c->own_stream); delete c; return rc; }
    rc = hip_check(hipEventCreateWithFlags(&c->aux_fork, hipEventDisableTiming), "hipEventCreate");
    if (!rc) rc = hip_check(hipEventCreateWithFlags(&c->aux_join, hipEventDisableTiming), "hipEventCreate");
    if (!rc) rc = hip_check(hipEventCreateWithFlags(&c->aux_mid, hipEventDisableTiming), "hipEventCreate");
    if (rc) { (void)hipStreamDestroy(c->aux_stream); (void)hipStreamDestroy(c->own_stream); delete c; return rc; }
    c->forked_from = nullptr;
    c->stream = c->own_stream;
    *out = c;
    return ENBL_OK;
}

int enbl_context_destroy(enbl_context* ctx)
{
    if (!ctx) return ENBL_OK;
    (void)hipSetDevice(ctx->device);
    (void)sync_streams(ctx);
    for (auto& c : ctx->cache) enbl_pyramid_destroy(c.p);
    batch_release(ctx);
    for (auto& t : ctx->timed) { (void)hipEventDestroy(t.start); (void)hipEventDestroy(t.stop); }
    if (ctx->scratch_alpha) (void)hipFree(ctx->scratch_alpha);
    for (int i = 0; i < kScratchSlots; ++i) if (ctx->scratch_buf[i]) (void)hipFree(ctx->scratch_buf[i]);
    (void)hipEventDestroy(ctx->aux_fork);
    (void)hipEventDestroy(ctx->aux_join);
    (void)hipEventDestroy(ctx->aux_mid);
    (void)hipStreamDestroy(ctx->aux_stream);
    (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return ENBL_OK;
}

int enbl_context_set_stream(enbl_context* ctx, void* s)
{
    if (!ctx) return fail(ENBL_E_ARG, "null context");
    ctx->stream = s ? (hipStream_t)s : ctx->own_stream;
    return ENBL_OK;
}

int enbl_context_get_stream(const enbl_context* ctx, void** s)
{
    if (!ctx || !s) return fail(ENBL_E_ARG, "null argument");
    *s = (void*)ctx->stream;
    return ENBL_OK;
}

int enbl_device_alloc(enbl_context* ctx, size_t bytes, void** out)
{
    if (!ctx || !out) return fail(ENBL_E_ARG, "enbl_device_alloc: null argument");
    *out = nullptr;
    if (bytes == 0) return ENBL_OK;
    int rc = hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    if (rc) return rc;
    return hip_check(hipMalloc(out, bytes), "enbl_device_alloc");
}

int enbl_device_free(enbl_context* ctx, void* p)
{
    if (!ctx) return fail(ENBL_E_ARG, "enbl_device_free: null context");
    if (!p) return ENBL_OK;
    int rc = hip_check(hipStreamSynchronize(ctx->stream), "enbl_device_free");
    if (rc) return rc;
    return hip_check(hipFree(p), "enbl_device_free");
}

int enbl_copy(enbl_context* ctx, void* dst, const void* src, size_t bytes)
{
    if (!ctx || (!dst && bytes) || (!src && bytes)) return fail(ENBL_E_ARG, "enbl_copy: null argument");
    if (bytes == 0) return ENBL_OK;
    int rc = hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream), "enbl_copy");
    if (rc) return rc;
    return hip_check(hipStreamSynchronize(ctx->stream), "enbl_copy");
}

int enbl_set_kernel_policy(int policy)
{
    if (policy < ENBL_KERNELS_AUTO || policy > ENBL_KERNELS_STREAMING) return fail(ENBL_E_ARG, "kernel policy %d", policy);
    g_policy = policy;
    return ENBL_OK;
}

int enbl_synchronize(enbl_context* ctx)
{
    if (!ctx) return fail(ENBL_E_ARG, "null context");
    return hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

int enbl_profile_enable(enbl_context* ctx, int on)
{
    if (!ctx) return fail(ENBL_E_ARG, "null context");
    ctx->profile = on != 0;
    return ENBL_OK;
}

int enbl_profile_report(enbl_context* ctx, char* buf, size_t n)
{
    if (!ctx || !buf || n == 0) return fail(ENBL_E_ARG, "null argument");
    int rc = hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (rc) return rc;
    struct Acc { double ms = 0.0, bytes = 0.0, own = 0.0; int n = 0; };
    std::map<std::string, Acc> acc;
    for (auto& t : ctx->timed) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
            auto& e = acc[t.name];
            e.ms += ms;
            e.bytes += t.bytes;
            e.own += t.own;
            e.n += 1;
        }
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    ctx->timed.clear();
    std::string out;
    for (auto& kv : acc) {
        char line[200];
        snprintf(line, sizeof line, "%s %.6f %d %.0f %.0f\n", kv.first.c_str(), kv.second.ms, kv.second.n,
                 kv.second.bytes, kv.second.own);
        out += line;
    }
    snprintf(buf, n, "%s", out.c_str());
    return ENBL_OK;
}

int enbl_reduce(enbl_context* ctx, int ptype, int C, int wrap,
                const void* src, int sw, int sh, long sp, long spl, const uint8_t* sa, long sap,
                void* dst, int dw, int dh, long dp, long dpl, uint8_t* da, long dap)
{
    if (!ctx || !src || !dst) return fail(ENBL_E_ARG, "enbl_reduce: null argument");
    if (sa && !da) return fail(ENBL_E_ARG, "enbl_reduce: alpha source needs a destination alpha");
    if (ptype_size(ptype) == 0 || ptype > ENBL_F64) return fail(ENBL_E_UNSUPPORTED, "enbl_reduce: pyramid type %d", ptype);
    if ((sp & 1) || (spl & 1) || !aligned2(src, ptype))
        return fail(ENBL_E_ARG, "enbl_reduce: source rows must start on even elements");
    return do_reduce(ctx, ptype, C, wrap, src, sw, sh, sp, spl, sa, sap, dst, dw, dh, dp, dpl, da, dap);
}

int enbl_expand(enbl_context* ctx, int ptype, int C, int add, int wrap,
                const void* src, int sw, int sh, long sp, long spl,
                void* dst, int dw, int dh, long dp, long dpl)
{
    if (!ctx || !src || !dst) return fail(ENBL_E_ARG, "enbl_expand: null argument");
    if (ptype_size(ptype) == 0 || ptype > ENBL_F64) return fail(ENBL_E_UNSUPPORTED, "enbl_expand: pyramid type %d", ptype);
    if ((dp & 1) || (dpl & 1) || !aligned2(dst, ptype))
        return fail(ENBL_E_ARG, "enbl_expand: destination rows must start on even elements");
    return do_expand(ctx, ptype, C, add, wrap, src, sw, sh, sp, spl, dst, dw, dh, dp, dpl);
}

int enbl_pyramid_create(enbl_context* ctx, int ptype, int C, int w, int h, int levels, enbl_pyramid** out)
{
    if (!ctx || !out) return fail(ENBL_E_ARG, "enbl_pyramid_create: null argument");
    if (ptype_size(ptype) == 0 || ptype > ENBL_F64) return fail(ENBL_E_UNSUPPORTED, "pyramid type %d", ptype);
    if (C != 1 && C != 3) return fail(ENBL_E_UNSUPPORTED, "pyramid channels %d", C);
    if (w < 1 || h < 1) return fail(ENBL_E_SIZE, "empty pyramid");
    if (levels < 1 || levels > 29) return fail(ENBL_E_PRECONDITION, "filterHalfWidth: levels outside of permissible range");
    enbl_pyramid* p = new enbl_pyramid();
    p->ctx = ctx;
    p->ptype = ptype; p->channels = C; p->w = w; p->h = h; p->levels = levels;
    p->lv = plan_levels(ptype, C, w, h, levels, &p->bytes);
    p->base = nullptr;
    int rc = hip_check(hipMalloc(&p->base, p->bytes), "pyramid allocation");
    if (rc) { delete p; return rc; }
    *out = p;
    return ENBL_OK;
}

int enbl_pyramid_destroy(enbl_pyramid* p)
{
    if (!p) return ENBL_OK;
    if (p->base) (void)hipFree(p->base);
    delete p;
    return ENBL_OK;
}

int enbl_pyramid_levels(const enbl_pyramid* p) { return p ? p->levels : 0; }

int enbl_pyramid_level(const enbl_pyramid* p, int l, void** data, int* w, int* h, long* pitch, long* plane)
{
    if (!p || l < 0 || l >= p->levels) return fail(ENBL_E_ARG, "enbl_pyramid_level: bad level");
    if (data) *data = level_ptr(p, l);
    if (w) *w = p->lv[l].w;
    if (h) *h = p->lv[l].h;
    if (pitch) *pitch = p->lv[l].pitch;
    if (plane) *plane = p->lv[l].plane;
    return ENBL_OK;
}

int enbl_pyramid_copy_level(enbl_pyramid* p, int l, void* buf, int to_pyramid)
{
    if (!p || !buf || l < 0 || l >= p->levels) return fail(ENBL_E_ARG, "enbl_pyramid_copy_level: bad argument");
    const Level& L = p->lv[l];
    const size_t es = (size_t)ptype_size(p->ptype);
    hipStream_t s = p->ctx->stream;
    for (int c = 0; c < p->channels; ++c) {
        char* lvl = (char*)level_ptr(p, l) + (size_t)c * L.plane * es;
        char* pk = (char*)buf + (size_t)c * L.w * L.h * es;
        hipError_t e = to_pyramid
            ? hipMemcpy2DAsync(lvl, L.pitch * es, pk, L.w * es, L.w * es, L.h, hipMemcpyDefault, s)
            : hipMemcpy2DAsync(pk, L.w * es, lvl, L.pitch * es, L.w * es, L.h, hipMemcpyDefault, s);
        int rc = hip_check(e, "enbl_pyramid_copy_level");
        if (rc) return rc;
    }
    return hip_check(hipStreamSynchronize(s), "enbl_pyramid_copy_level");
}

int enbl_auto_levels(int w, int h, int minimum_levels, int exact_levels)
{
    unsigned s = (unsigned)std::max(0, std::min(w, h));
    const unsigned lo = minimum_levels < 1 ? 1u : (unsigned)minimum_levels;
    unsigned levels = lo;
    while (levels <= 29) {
        if (s <= 8u) break;
        s = (s + 1u) >> 1;
        ++levels;
    }
    if (levels > lo) {
        if (exact_levels >= 1) levels = std::min(levels, (unsigned)exact_levels);
        else if (exact_levels < 0) levels = ((int)levels + exact_levels >= (int)lo) ? levels - (unsigned)(-exact_levels) : lo;
    }
    return (int)std::min(levels, 29u);
}

int enbl_gaussian_pyramid(enbl_context* ctx, int wrap, int itype, int frac, const void* img, long ipitch,
                          const uint8_t* alpha, long apitch, enbl_pyramid* out)
{
    if (!ctx || !img || !out) return fail(ENBL_E_ARG, "enbl_gaussian_pyramid: null argument");
    int rc = build_level0(ctx, out, itype, frac, img, ipitch);
    if (rc) return rc;
    return gaussian_levels(ctx, out, wrap, alpha, apitch);
}

int enbl_copy_to_pyramid(enbl_context* ctx, int itype, int frac, const void* img, long ipitch, enbl_pyramid* out)
{
    if (!ctx || !img || !out) return fail(ENBL_E_ARG, "enbl_copy_to_pyramid: null argument");
    return build_level0(ctx, out, itype, frac, img, ipitch);
}

int enbl_laplacian_pyramid(enbl_context* ctx, int wrap, int itype, int frac, const void* img, long ipitch,
                           const uint8_t* alpha, long apitch, enbl_pyramid* out)
{
    if (!ctx || !img || !out) return fail(ENBL_E_ARG, "enbl_laplacian_pyramid: null argument");
    return laplacian_pyramid_any(ctx, wrap, itype, frac, img, ipitch, alpha, apitch, out);
}

int enbl_pyramid_build_levels(enbl_context* ctx, enbl_pyramid* p, int wrap, const uint8_t* alpha, long apitch,
                              int laplacian)
{
    if (!ctx || !p) return fail(ENBL_E_ARG, "enbl_pyramid_build_levels: null argument");
    int rc = gaussian_levels(ctx, p, wrap, alpha, apitch);
    if (!rc && laplacian) rc = laplacian_levels(ctx, p, wrap);
    return rc;
}

int enbl_collapse_pyramid(enbl_context* ctx, int wrap, enbl_pyramid* p)
{
    if (!ctx || !p) return fail(ENBL_E_ARG, "enbl_collapse_pyramid: null argument");
    return collapse_levels(ctx, wrap, p, 0);
}

int enbl_blend(enbl_context* ctx, const enbl_pyramid* m, const enbl_pyramid* wlp, enbl_pyramid* blp, double wv)
{
    if (!ctx || !m || !wlp || !blp) return fail(ENBL_E_ARG, "enbl_blend: null argument");
    if (m->channels != 1 || wlp->channels != blp->channels || wlp->ptype != blp->ptype ||
        m->levels < blp->levels || wlp->levels < blp->levels)
        return fail(ENBL_E_SIZE, "enbl_blend: pyramid shapes disagree");
    for (int l = 0; l < blp->levels; ++l) {
        const Level& L = blp->lv[l];
        if (m->lv[l].w != L.w || m->lv[l].h != L.h || wlp->lv[l].w != L.w || wlp->lv[l].h != L.h)
            return fail(ENBL_E_SIZE, "enbl_blend: level %d sizes disagree", l);
        Timed t(ctx, "blend", (double)L.w * L.h * (ptype_size(m->ptype) + 3.0 * blp->channels * ptype_size(blp->ptype)));
        int rc = launch_blend(m->ptype, blp->ptype, blp->channels, L.w, L.h, level_ptr(m, l), m->lv[l].pitch,
                              level_ptr(wlp, l), wlp->lv[l].pitch, wlp->lv[l].plane,
                              level_ptr(blp, l), L.pitch, L.plane, wv, ctx->stream);
        if (rc) return fail(rc, "enbl_blend: unsupported mask/pyramid types");
        rc = launch_status("blend");
        if (rc) return rc;
    }
    return ENBL_OK;
}

int enbl_copy_from_pyramid_if(enbl_context* ctx, const enbl_pyramid* p, int level, int frac,
                              const uint8_t* mask, long mpitch, void* img, int itype, long ipitch,
                              int dither, uint64_t seed)
{
    if (!ctx || !p || !mask || !img) return fail(ENBL_E_ARG, "enbl_copy_from_pyramid_if: null argument");
    if (level < 0 || level >= p->levels) return fail(ENBL_E_ARG, "bad level");
    const Level& L = p->lv[level];
    Timed t(ctx, "from_pyramid", (double)L.w * L.h * (1.0 + p->channels * (ptype_size(p->ptype) + ptype_size(itype))));
    int rc = launch_from_pyramid_if(p->ptype, itype, frac, p->channels, level_ptr(p, level), L.pitch, L.plane,
                                    L.w, L.h, mask, mpitch, img, ipitch, dither, seed, ctx->stream);
    if (rc) return fail(rc, "copyFromPyramidImageIf: unsupported pyramid/image type pair");
    return launch_status("from_pyramid");
}

namespace {
int enblend_pair_rgb(enbl_context* ctx, int itype, int C, int w, int h, int levels, int wrap,
                     const void* white, long wpitch, const uint8_t* walpha, long wapitch,
                     void* black, long bpitch, uint8_t* balpha, long bapitch,
                     const uint8_t* mask, long mpitch, int dither, uint64_t seed);

// The step in the CIECAM02 blend space: the JCh level 0 carries fraction
// bits, so it is written straight into each image's pyramid
// (copyToPyramidImage with ConvertVectorToJCHPyramidFunctor), the levels
// follow level by level, and the collapsed level 0 goes back to RGB under
// the union alpha (copyFromPyramidImageIf with
// ConvertJCHPyramidToVectorFunctor; no dither: the pyramid is read as
// double).  src/enblend.h:378-510, src/fixmath.h:2018-2294.
int enblend_pair_jch(enbl_context* ctx, int itype, int w, int h, int levels, int wrap,
                     const void* white, long wpitch, const uint8_t* walpha, long wapitch,
                     void* black, long bpitch, uint8_t* balpha, long bapitch, const uint8_t* mask, long mpitch)
{
    const int pt = image_ptype_for(itype), mt = mask_ptype_for(itype);
    const int mfrac = itype == ENBL_U8 ? 7 : 15;
    enbl_pyramid *mgp = nullptr, *wlp = nullptr, *blp = nullptr;
    int rc = cached_pyramid(ctx, kRoleMask, mt, 1, w, h, levels, 0, &mgp);
    if (!rc) rc = cached_pyramid(ctx, kRoleImage, pt, 3, w, h, levels, 0, &wlp);
    if (!rc) rc = cached_pyramid(ctx, kRoleImage, pt, 3, w, h, levels, 1, &blp);
    if (rc) return rc;
    rc = enbl_gaussian_pyramid(ctx, wrap, ENBL_U8, mfrac, mask, mpitch, nullptr, 0, mgp);
    if (rc) return rc;
    const void* img[2] = {white, black};
    const long ip[2] = {wpitch, bpitch};
    const uint8_t* al[2] = {walpha, balpha};
    const long ap[2] = {wapitch, bapitch};
    enbl_pyramid* lp[2] = {wlp, blp};
    for (int k = 0; k < 2; ++k) {
        const Level& L0 = lp[k]->lv[0];
        rc = rgb_to_jch(ctx, itype, w, h, img[k], ip[k], level_ptr(lp[k], 0), L0.pitch, L0.plane);
        if (!rc) rc = gaussian_levels(ctx, lp[k], wrap, al[k], ap[k]);
        if (!rc) rc = laplacian_levels(ctx, lp[k], wrap);
        if (rc) return rc;
    }
    {
        Timed t(ctx, "alpha_union", 3.0 * w * h);
        launch_alpha_union(walpha, wapitch, balpha, bapitch, w, h, ctx->stream);
        rc = launch_status("alpha_union");
        if (rc) return rc;
    }
    rc = enbl_blend(ctx, mgp, wlp, blp, (double)(255 << mfrac));
    if (!rc) rc = enbl_collapse_pyramid(ctx, wrap, blp);
    if (rc) return rc;
    const Level& B0 = blp->lv[0];
    return jch_to_rgb_if(ctx, itype, w, h, level_ptr(blp, 0), B0.pitch, B0.plane, balpha, bapitch, black, bpitch);
}
}  // namespace

// One blend step; in a CIELAB / CIELUV blend space both RGB images are
// converted into scaled blend-space images first, the step runs on those,
// and its result goes back to RGB under the union alpha, as the reference's
// converters around copyToPyramidImage / copyFromPyramidImageIf do
// (src/fixmath.h:2018-2294); CIECAM02 takes enblend_pair_jch.
int enbl_enblend_pair(enbl_context* ctx, int itype, int C, int w, int h, int levels, int wrap,
                      const void* white, long wpitch, const uint8_t* walpha, long wapitch,
                      void* black, long bpitch, uint8_t* balpha, long bapitch,
                      const uint8_t* mask, long mpitch, int dither, uint64_t seed)
{
    if (!ctx || !white || !walpha || !black || !balpha || !mask) return fail(ENBL_E_ARG, "enbl_enblend_pair: null argument");
    const int cs = C == 3 ? ctx->colorspace : ENBL_COLORSPACE_IDENTITY;
    if (cs == ENBL_COLORSPACE_IDENTITY)
        return enblend_pair_rgb(ctx, itype, C, w, h, levels, wrap, white, wpitch, walpha, wapitch, black, bpitch,
                                balpha, bapitch, mask, mpitch, dither, seed);
    int rc = check_blend_space(cs, itype, C);
    if (rc) return rc;
    if (w < 1 || h < 1) return fail(ENBL_E_SIZE, "enbl_enblend_pair: empty image");
    if (cs == ENBL_COLORSPACE_CIECAM)
        return enblend_pair_jch(ctx, itype, w, h, levels, wrap, white, wpitch, walpha, wapitch, black, bpitch, balpha,
                                bapitch, mask, mpitch);
    const size_t es = itype == ENBL_U8 ? 1 : 2, plane = (size_t)w * h * 3 * es;
    void* buf;
    rc = scratch(ctx, 32, 2 * plane, &buf);
    if (rc) return rc;
    char* wl = (char*)buf;
    char* bl = wl + plane;
    rc = rgb_to_blend_space(ctx, cs, itype, w, h, white, wpitch, wl, 3L * w);
    if (!rc) rc = rgb_to_blend_space(ctx, cs, itype, w, h, black, bpitch, bl, 3L * w);
    if (!rc) rc = enblend_pair_rgb(ctx, itype, C, w, h, levels, wrap, wl, 3L * w, walpha, wapitch, bl, 3L * w,
                                   balpha, bapitch, mask, mpitch, dither, seed);
    if (!rc) rc = blend_space_to_rgb_if(ctx, cs, itype, w, h, bl, 3L * w, balpha, bapitch, black, bpitch);
    return rc;
}

namespace {
int enblend_pair_rgb(enbl_context* ctx, int itype, int C, int w, int h, int levels, int wrap,
                     const void* white, long wpitch, const uint8_t* walpha, long wapitch,
                     void* black, long bpitch, uint8_t* balpha, long bapitch,
                     const uint8_t* mask, long mpitch, int dither, uint64_t seed)
{
    if (itype != ENBL_U8 && itype != ENBL_U16 && itype != ENBL_F32) return fail(ENBL_E_UNSUPPORTED, "image type %d", itype);
    const int pt = image_ptype_for(itype);
    const int mt = mask_ptype_for(itype);
    const int ifrac = itype == ENBL_F32 ? 0 : 7;
    const int mfrac = itype == ENBL_U8 ? 7 : 15;
    enbl_pyramid *mgp = nullptr, *wlp = nullptr, *blp = nullptr;
    int rc = cached_pyramid(ctx, kRoleMask, mt, 1, w, h, levels, 0, &mgp);
    if (!rc) rc = cached_pyramid(ctx, kRoleImage, pt, C, w, h, levels, 0, &wlp);
    if (!rc) rc = cached_pyramid(ctx, kRoleImage, pt, C, w, h, levels, 1, &blp);
    if (rc) return rc;

    const double wv = (double)(255 << mfrac);
    const int s = fused_level_count(wlp, wrap, walpha);
    if (s == 0) {
        // Mask Gaussian pyramid without alpha (src/enblend.h:384-388).
        rc = enbl_gaussian_pyramid(ctx, wrap, ENBL_U8, mfrac, mask, mpitch, nullptr, 0, mgp);
        if (rc) return rc;
        // Laplacian pyramids of white and black (src/enblend.h:419-446).
        rc = enbl_laplacian_pyramid(ctx, wrap, itype, ifrac, white, wpitch, walpha, wapitch, wlp);
        if (rc) return rc;
        rc = enbl_laplacian_pyramid(ctx, wrap, itype, ifrac, black, bpitch, balpha, bapitch, blp);
        if (rc) return rc;
        // Union alpha (src/enblend.h:461-463).
        {
            Timed t(ctx, "alpha_union", 3.0 * w * h);
            launch_alpha_union(walpha, wapitch, balpha, bapitch, w, h, ctx->stream);
            rc = launch_status("alpha_union");
            if (rc) return rc;
        }
        // blend with white = 255 << mask fraction bits (src/enblend.h:472-474).
        rc = enbl_blend(ctx, mgp, wlp, blp, wv);
        if (rc) return rc;
        rc = enbl_collapse_pyramid(ctx, wrap, blp);
        if (rc) return rc;
        return enbl_copy_from_pyramid_if(ctx, blp, 0, ifrac, balpha, bapitch, black, itype, bpitch, dither, seed);
    }
    // Mask Gaussian pyramid without alpha (src/enblend.h:384-388): the large
    // levels up to s here, the rest with the image tails.  Level 1 is reduced
    // straight from the mask bytes, converted on the fly (the fixed-point
    // level 0 is not stored: the level-0 collapse reads the bytes).
    {
        const Level& D = mgp->lv[1];
        if (w < 2 || h < 2) return fail(ENBL_E_PRECONDITION, "src image too small in reduce");
        ReduceArgs a;
        memset(&a, 0, sizeof a);
        a.src = mask; a.spitch = mpitch; a.splane = 0; a.sw = w; a.sh = h;
        a.dst = level_ptr(mgp, 1); a.dpitch = D.pitch; a.dplane = D.plane; a.dw = D.w; a.dh = D.h;
        a.channels = 1; a.frac = mfrac;
        char name[64];
        snprintf(name, sizeof name, "reduce_mask0_%s", pname(mt));
        const double esm = ptype_size(mt);
        // credited bytes: the conversion of level 0 and the reduce it feeds;
        // own bytes: the mask bytes read once, level 1 written once
        Timed t(ctx, name, (double)w * h * (1.0 + esm) + esm * ((double)w * h + (double)D.w * D.h),
                (double)w * h + esm * (double)D.w * D.h);
        rc = launch_reduce_image(ENBL_U8, mt, a, ctx->stream);
        if (rc) return fail(rc, "mask reduce: unsupported type");
        rc = launch_status(name);
        if (rc) return rc;
    }
    for (int l = 2; !rc && l <= s; ++l) {
        const Level& S_ = mgp->lv[l - 1];
        const Level& D = mgp->lv[l];
        rc = do_reduce(ctx, mt, 1, 0, level_ptr(mgp, l - 1), S_.w, S_.h, S_.pitch, S_.plane, nullptr, 0,
                       level_ptr(mgp, l), D.w, D.h, D.pitch, D.plane, nullptr, 0);
    }
    if (rc) return rc;
    // Laplacian pyramids of white and black (src/enblend.h:419-446): the
    // fused levels, each pyramid keeping its own alpha of level s for the
    // shared small-level launches.
    const int esP = ptype_size(pt), esM = ptype_size(mt);
    const bool otf = s >= 2 && itype != ENBL_F32 && C == 3;
    void* g1[2] = {nullptr, nullptr};
    {
        uint8_t* wab[2];
        uint8_t* bab[2];
        long apl;
        void* bp;
        rc = alpha_pingpong(ctx, wlp, &wab[0], &wab[1], &apl);
        if (!rc) rc = scratch(ctx, 14, 2 * (size_t)apl * wlp->lv[1].h, &bp);
        if (rc) return rc;
        bab[0] = (uint8_t*)bp;
        bab[1] = (uint8_t*)bp + (size_t)apl * wlp->lv[1].h;
        if (otf) {
            // Level 0 of either Laplacian pyramid is never stored: only the
            // Gaussian levels 1 (and their alpha) are reduced from the
            // images here, kept apart for the level-0 collapse, which forms
            // both Laplacian levels 0 on the fly.
            const Level& D = wlp->lv[1];
            const size_t gbytes = round_up((size_t)D.plane * C * esP, 256);
            void* g;
            rc = scratch(ctx, 21, 2 * gbytes, &g);
            if (rc) return rc;
            g1[0] = g;
            g1[1] = (char*)g + gbytes;
            const void* src[2] = {white, black};
            const long sp[2] = {wpitch, bpitch};
            const uint8_t* sa[2] = {walpha, balpha};
            const long sap[2] = {wapitch, bapitch};
            uint8_t* da[2] = {wab[0], bab[0]};
            for (int k = 0; k < 2 && !rc; ++k) {
                ReduceArgs r;
                memset(&r, 0, sizeof r);
                r.src = src[k]; r.spitch = sp[k]; r.splane = 0; r.sw = w; r.sh = h;
                r.salpha = sa[k]; r.sapitch = sap[k];
                r.dst = g1[k]; r.dpitch = D.pitch; r.dplane = D.plane; r.dw = D.w; r.dh = D.h;
                r.dalpha = da[k]; r.dapitch = apl;
                r.channels = C; r.frac = ifrac;
                char name[64];
                snprintf(name, sizeof name, "reduce_image0_%s_c%d", pname(pt), C);
                const double n0 = (double)w * h, n1 = (double)D.w * D.h, esC = (double)esP * C;
                // credited bytes: the level-0 conversion and the reduce it
                // feeds; own bytes: image and alpha read, level 1 and its alpha
                // written
                Timed t(ctx, name, n0 * C * (ptype_size(itype) + (double)esP) + (esC + 1.0) * (n0 + n1),
                        n0 * (C * (double)ptype_size(itype) + 1.0) + n1 * (esC + 1.0));
                rc = launch_reduce_image(itype, pt, r, ctx->stream);
                if (rc) return fail(rc, "level-0 reduce: unsupported type");
                rc = launch_status(name);
            }
            if (!rc) rc = fused_laplacian_levels(ctx, itype, ifrac, white, wpitch, walpha, wapitch, wlp, s, wab, apl, 1, g1[0]);
            if (!rc) rc = fused_laplacian_levels(ctx, itype, ifrac, black, bpitch, balpha, bapitch, blp, s, bab, apl, 1, g1[1]);
        } else {
            rc = fused_laplacian_levels(ctx, itype, ifrac, white, wpitch, walpha, wapitch, wlp, s, wab, apl);
            if (!rc) rc = fused_laplacian_levels(ctx, itype, ifrac, black, bpitch, balpha, bapitch, blp, s, bab, apl);
        }
        if (!rc) rc = enblend_tail(ctx, mgp, wlp, blp, s, wab[(s - 1) & 1], bab[(s - 1) & 1], apl, wv);
        if (rc) return rc;
    }
    // Levels s-1..0: blend, add the expanded coarser level and (level 0)
    // convert in one pass.
    for (int l = s - 1; l >= 0; --l) {
        const Level& L = blp->lv[l];
        const Level& Cl = blp->lv[l + 1];
        FusedCollapseArgs a;
        memset(&a, 0, sizeof a);
        if (l == 0 && otf) {
            a.otf = 1;
            a.wimg = white; a.wipitch = wpitch;
            a.wg1 = g1[0]; a.wg1pitch = Cl.pitch; a.wg1plane = Cl.plane;
            a.bg1 = g1[1]; a.bg1pitch = Cl.pitch; a.bg1plane = Cl.plane;
        }
        a.coarse = level_ptr(blp, l + 1); a.cpitch = Cl.pitch; a.cplane = Cl.plane; a.cw = Cl.w; a.ch = Cl.h;
        a.black = level_ptr(blp, l); a.fpitch = L.pitch; a.fplane = L.plane;
        a.white = level_ptr(wlp, l); a.wpitch = wlp->lv[l].pitch; a.wplane = wlp->lv[l].plane;
        a.fw = L.w; a.fh = L.h;
        a.white_value = wv;
        a.img = black; a.ipitch = bpitch; a.ifrac = ifrac;
        a.walpha = walpha; a.wapitch = wapitch; a.balpha = balpha; a.bapitch = bapitch;
        a.dither = dither; a.seed = seed; a.rows = 0;
        if (l == 0) { a.mask = mask; a.mpitch = mpitch; a.mfrac = mfrac; }
        else { a.mask = level_ptr(mgp, l); a.mpitch = mgp->lv[l].pitch; a.mfrac = 0; }
        char name[64];
        snprintf(name, sizeof name, "collapse_level%s_%s_c%d", l == 0 ? "0" : "", pname(pt), C);
        // credited bytes: blend + expand-add (+ masked conversion and alpha
        // union at level 0); own bytes: the coarse level, both Laplacian
        // levels and the mask read once, the result written once (level 0:
        // the two alphas read, the image and the union alpha written)
        const double nL = (double)L.w * L.h, nC = (double)Cl.w * Cl.h;
        double bytes = nL * (esM + 3.0 * C * esP) + C * esP * (nC + 2.0 * nL);
        if (l == 0) bytes += nL * (1.0 + C * (esP + ptype_size(itype))) + 3.0 * w * h;
        double own = C * esP * (nC + 2.0 * nL) +
                     (l == 0 ? nL * (1.0 + 2.0 + C * ptype_size(itype) + 1.0) : nL * (esM + C * esP));
        if (a.otf) {
            // also the two expand-subtracts of level 0 (with the images'
            // conversion); own: the images and both Gaussian levels 1 are read
            // in place of the two Laplacian levels 0
            bytes += 2.0 * (C * esP * (nC + 2.0 * nL) + nL * C * (ptype_size(itype) + (double)esP));
            own = C * esP * 3.0 * nC + nL * (1.0 + 2.0 + 2.0 * C * ptype_size(itype)) + nL * (C * ptype_size(itype) + 1.0);
        }
        Timed t(ctx, name, bytes, own);
        rc = launch_collapse_level(pt, mt, l == 0 ? itype : 0, l == 0, a, C, ctx->stream);
        if (rc) return fail(rc, "collapse level: unsupported types");
        rc = launch_status(name);
        if (rc) return rc;
    }
    return ENBL_OK;
}
}  // namespace

}  // extern "C"
