// C ABI of batched enfuse (declared in include/enblend_amd.h):
// enbl_enfuse_batch, the launch-shape tuning and its autotuner.
//
// A batch is planned on the host into one arena of device pyramids (per
// stack: the image and alpha pyramids of every input above level 0, the
// mask pyramid of every input, the union-alpha pyramid and the result
// levels above level 0) and one table of job descriptors; the passes of
// batch.hip then run as 1 + (L-1) + L launches for the deepest stack's L
// levels, whatever the number of stacks.  The launch sequence is captured
// into a HIP graph and replayed while the batch (geometry, pointers,
// parameters, launch shapes) stays the same.
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

struct BatchState {
    void* arena = nullptr;
    size_t arena_bytes = 0;
    void* tables = nullptr;
    size_t table_bytes = 0;
    std::vector<unsigned char> key;   // uploaded table image + launch list behind `exec`
    hipGraphExec_t exec = nullptr;
    hipStream_t exec_stream = nullptr;
    int cus = 0;
    // exposure table cache (built-in functions only)
    std::vector<float> lut;
    int lut_itype = -1;
    double lut_key[4] = {0, 0, 0, 0};
    int lut_fn = -1;
};

namespace host {

void batch_release(enbl_context* ctx)
{
    BatchState* b = ctx->batch;
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->arena) (void)hipFree(b->arena);
    if (b->tables) (void)hipFree(b->tables);
    delete b;
    ctx->batch = nullptr;
}

}  // namespace host
}  // namespace enbl

namespace {

const int kDefaultRows[kBatchKinds] = {16, 16, 8};
const int kDefaultBpc[kBatchKinds] = {8, 8, 8};
const char* const kKindName[kBatchKinds] = {"batch_weights", "batch_reduce", "batch_fuse"};
const int kStrip = 62;   // owned columns of a wave strip (reduce_stream.h, batch.hip)

struct Tuning {
    int rows[kBatchKinds];
    int bpc[kBatchKinds];
};

Tuning context_tuning(const enbl_context* ctx)
{
    Tuning t;
    for (int k = 0; k < kBatchKinds; ++k) {
        t.rows[k] = ctx->batch_rows[k] > 0 ? ctx->batch_rows[k] : kDefaultRows[k];
        t.bpc[k] = ctx->batch_bpc[k] > 0 ? ctx->batch_bpc[k] : kDefaultBpc[k];
    }
    return t;
}

struct Launch {
    int kind;
    int level0;
    size_t jobs, tile0;   // table offsets
    int njobs, total;
    int rows, grid;
    double bytes;         // algorithmic bytes of the launch
};

// Everything one batch run needs besides the device buffers.
struct Plan {
    std::vector<unsigned char> tab;   // host image of the job tables
    std::vector<Launch> launches;
    size_t lut_off = 0;
    double w_exp = 0, w_sat = 0, scale = 0, maxpyr = 0;
    int frac = 0, dither = 0;
    int max_k = 0;   // deepest stack
    size_t images_off = 0, planes_off = 0;   // BImage / BPlane tables
    bool stack_weights = false;   // weights made stack by stack (enfuse_stack_masks), no weights launch
    // per stack: its images' mask level 0 and its union alpha in the arena
    std::vector<std::vector<void*>> mask0;
    std::vector<long> mpitch, upitch;
    std::vector<uint8_t*> ualpha;
};

// Device layout of one stack inside the arena.
struct StackLayout {
    int L;
    std::vector<Level> img;   // image pyramid levels 1..L-1 (index l-1), P x 3
    std::vector<Level> ial;   // image alpha levels 1..L-1, u8
    std::vector<Level> msk;   // mask pyramid levels 0..L-1, M
    std::vector<Level> ual;   // union alpha levels 0..L-1, u8
    std::vector<Level> res;   // result levels 1..L-1, P x 3
    size_t img_bytes = 0, ial_bytes = 0, msk_bytes = 0, ual_bytes = 0, res_bytes = 0;
    size_t base = 0;
    size_t img_at(int m) const { return base + (size_t)m * img_bytes; }
    size_t ial_at(int K, int m) const { return base + (size_t)K * img_bytes + (size_t)m * ial_bytes; }
    size_t msk_at(int K, int m) const { return base + (size_t)K * (img_bytes + ial_bytes) + (size_t)m * msk_bytes; }
    size_t ual_at(int K) const { return base + (size_t)K * (img_bytes + ial_bytes + msk_bytes); }
    size_t res_at(int K) const { return ual_at(K) + ual_bytes; }
    size_t end(int K) const { return res_at(K) + res_bytes; }
};

struct Table {
    std::vector<unsigned char>& buf;
    char* dbase;   // device address of the table buffer (null while sizing)
    explicit Table(std::vector<unsigned char>& b, char* d) : buf(b), dbase(d) {}
    template <class T> size_t put(const T* data, size_t n) {
        const size_t off = round_up(buf.size(), 64);
        buf.resize(off + sizeof(T) * n);
        if (n) memcpy(buf.data() + off, data, sizeof(T) * n);
        return off;
    }
    template <class T> const T* dev(size_t off) const { return reinterpret_cast<const T*>((uintptr_t)dbase + off); }
};

// The batched weights pass: exposure (average projector, no cutoff) and
// saturation of integer images, soft masks.
int batchable(int itype, const enbl_enfuse_params* p, int wrap)
{
    return (itype == ENBL_U8 || itype == ENBL_U16) && p->w_contrast <= 0.0 && p->w_entropy <= 0.0 && !p->hard_mask &&
           !p->exposure_cutoff && p->gray_projector == ENBL_GRAY_AVERAGE && !wrap;
}

// The batched pyramid passes: 8/16-bit and float images without wraparound,
// any weights (made stack by stack by the weight kernels of enbl_enfuse when
// the batched weights pass does not cover them).
int batchable_pyramids(int itype, int wrap)
{
    return (itype == ENBL_U8 || itype == ENBL_U16 || itype == ENBL_F32) && !wrap;
}

int check_stacks(int itype, int n, const enbl_stack* st, const enbl_enfuse_params* p)
{
    if (n < 0 || (n > 0 && !st)) return fail(ENBL_E_ARG, "enbl_enfuse_batch: bad stack list");
    for (int s = 0; s < n; ++s) {
        const enbl_stack& S = st[s];
        if (S.w < 1 || S.h < 1) return fail(ENBL_E_SIZE, "enbl_enfuse_batch: stack %d has size %dx%d", s, S.w, S.h);
        if (S.K < 1) return fail(ENBL_E_ARG, "enbl_enfuse_batch: stack %d has no images", s);
        if (S.levels < 1 || S.levels > 29)
            return fail(ENBL_E_PRECONDITION, "filterHalfWidth: levels outside of permissible range (stack %d)", s);
        if (!S.imgs || !S.ipitches || !S.alphas || !S.apitches || !S.out || !S.out_alpha)
            return fail(ENBL_E_ARG, "enbl_enfuse_batch: null pointer in stack %d", s);
        for (int m = 0; m < S.K; ++m)
            if (!S.imgs[m] || !S.alphas[m]) return fail(ENBL_E_ARG, "enbl_enfuse_batch: null image %d of stack %d", m, s);
        // reduce()'s precondition on every level that is reduced (src/pyramid.h:193-194).
        int w = S.w, h = S.h;
        for (int l = 0; l + 1 < S.levels; ++l) {
            if (w < 2 || h < 2) return fail(ENBL_E_PRECONDITION, "src image too small in reduce (stack %d)", s);
            w = (w + 1) >> 1;
            h = (h + 1) >> 1;
        }
        const int rc = check_params(itype, S.w, S.h, p);
        if (rc) return rc;
    }
    return ENBL_OK;
}

int state(enbl_context* ctx, BatchState** out)
{
    if (!ctx->batch) {
        BatchState* b = new BatchState();
        int rc = hip_check(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, ctx->device), "CU count");
        if (rc) { delete b; return rc; }
        ctx->batch = b;
    }
    *out = ctx->batch;
    return ENBL_OK;
}

void drop_graph(BatchState* b)
{
    if (b->exec) {
        (void)hipStreamSynchronize(b->exec_stream);
        (void)hipGraphExecDestroy(b->exec);
    }
    b->exec = nullptr;
    b->key.clear();
}

// Grows a device buffer of the batch state; the captured graph refers to the
// old one, so it goes too.
int grow(enbl_context* ctx, BatchState* b, void** buf, size_t* have, size_t need, const char* what)
{
    if (*have >= need) return ENBL_OK;
    int rc = hip_check(hipStreamSynchronize(ctx->stream), what);
    if (rc) return rc;
    drop_graph(b);
    if (*buf) (void)hipFree(*buf);
    *buf = nullptr;
    *have = 0;
    rc = hip_check(hipMalloc(buf, need), what);
    if (rc) return rc;
    *have = need;
    return ENBL_OK;
}

const std::vector<float>& lut_for(BatchState* b, int itype, const enbl_enfuse_params* p)
{
    const WeightArgs a = weight_args(p);
    const double k[4] = {p->w_exposure, a.exposure_optimum, a.exposure_width, 0.0};
    if (p->exposure_table || b->lut_itype != itype || b->lut_fn != a.exposure_function || memcmp(k, b->lut_key, sizeof k)) {
        exposure_lut(itype, p, a, b->lut);
        b->lut_itype = p->exposure_table ? -1 : itype;
        b->lut_fn = a.exposure_function;
        memcpy(b->lut_key, k, sizeof k);
    }
    return b->lut;
}

void layout_stacks(int itype, int n, const enbl_stack* st, std::vector<StackLayout>& lay, size_t* total)
{
    const int pt = image_ptype_for(itype), mt = mask_ptype_for(itype);
    size_t off = 0;
    lay.resize((size_t)n);
    for (int s = 0; s < n; ++s) {
        const enbl_stack& S = st[s];
        StackLayout& Y = lay[(size_t)s];
        Y.L = S.levels;
        const int w1 = (S.w + 1) >> 1, h1 = (S.h + 1) >> 1;
        if (S.levels > 1) {
            Y.img = plan_levels(pt, 3, w1, h1, S.levels - 1, &Y.img_bytes);
            Y.ial = plan_levels(ENBL_U8, 1, w1, h1, S.levels - 1, &Y.ial_bytes);
            Y.res = plan_levels(pt, 3, w1, h1, S.levels - 1, &Y.res_bytes);
        }
        Y.msk = plan_levels(mt, 1, S.w, S.h, S.levels, &Y.msk_bytes);
        Y.ual = plan_levels(ENBL_U8, 1, S.w, S.h, S.levels, &Y.ual_bytes);
        Y.base = off;
        off = Y.end(S.K);
    }
    *total = off;
}

// Builds the job tables and the launch list of a batch (dbase null: sizing).
void build_plan(int itype, int n, const enbl_stack* st, const std::vector<StackLayout>& lay, char* arena, char* dbase,
                const std::vector<float>& lut, const Tuning& tu, int cus, Plan& P)
{
    const int isz = ptype_size(itype);
    const int pt = image_ptype_for(itype), mt = mask_ptype_for(itype);
    const int es = ptype_size(pt), ms = ptype_size(mt);
    P.tab.clear();
    P.launches.clear();
    P.mask0.clear();
    P.mpitch.clear();
    P.upitch.clear();
    P.ualpha.clear();
    Table T(P.tab, dbase);
    P.lut_off = T.put(lut.data(), lut.size());
    auto tiles = [](int w, int h, int rows, int* tx) { *tx = (w + 63) / 64; return *tx * ((h + rows - 1) / rows); };
    auto add_launch = [&](int kind, int level0, size_t jobs, const std::vector<int>& t0, int total, double bytes) {
        Launch L{};
        L.kind = kind;
        L.level0 = level0;
        L.jobs = jobs;
        L.tile0 = T.put(t0.data(), t0.size());
        L.njobs = (int)t0.size();
        L.total = total;
        L.rows = tu.rows[kind];
        L.grid = std::max(1, std::min(total, cus * tu.bpc[kind]));
        L.bytes = bytes;
        P.launches.push_back(L);
    };
    auto at = [&](size_t off) { return arena + off; };
    std::vector<BImage> images;   // weights pass: per-image inputs
    std::vector<BPlane> planes;   // merge + collapse passes: per image fine, coarse, mask
    int Lmax = 0;
    P.max_k = 0;
    for (int s = 0; s < n; ++s) {
        Lmax = std::max(Lmax, st[s].levels);
        P.max_k = std::max(P.max_k, st[s].K);
    }

    // Weights.
    {
        std::vector<BWeightJob> jobs;
        std::vector<int> t0;
        int total = 0;
        double bytes = 0;
        for (int s = 0; s < n; ++s) {
            const enbl_stack& S = st[s];
            const StackLayout& Y = lay[(size_t)s];
            BWeightJob J{};
            J.w = S.w;
            J.h = S.h;
            J.K = S.K;
            J.tiles_x = (S.w + 127) / 128;
            const int nt = J.tiles_x * ((S.h + tu.rows[kBatchWeights] - 1) / tu.rows[kBatchWeights]);
            J.img0 = (int)images.size();
            for (int m = 0; m < S.K; ++m) {
                images.emplace_back();
                BImage& B = images.back();
                B.img = S.imgs[m];
                B.ipitch = S.ipitches[m];
                B.alpha = S.alphas[m];
                B.apitch = S.apitches[m];
                B.mask0 = at(Y.msk_at(S.K, m) + Y.msk[0].offset);
            }
            J.mpitch = Y.msk[0].pitch;
            J.ualpha = (uint8_t*)at(Y.ual_at(S.K) + Y.ual[0].offset);
            J.upitch = Y.ual[0].pitch;
            P.mask0.emplace_back();
            for (int m = 0; m < S.K; ++m) P.mask0.back().push_back(images[(size_t)J.img0 + m].mask0);
            P.mpitch.push_back(J.mpitch);
            P.ualpha.push_back(J.ualpha);
            P.upitch.push_back(J.upitch);
            J.fallback = (float)(255 / S.K);
            jobs.push_back(J);
            t0.push_back(total);
            total += nt;
            bytes += (double)S.w * S.h * ((double)S.K * (3.0 * isz + 1.0 + ms) + 1.0);
        }
        add_launch(kBatchWeights, 0, T.put(jobs.data(), jobs.size()), t0, total, bytes);
    }

    // Reduce passes: level l -> l + 1 of every image and mask pyramid.
    for (int l = 0; l + 1 < Lmax; ++l) {
        std::vector<BReduceJob> jobs;
        std::vector<int> t0;
        int total = 0;
        double bytes = 0;
        for (int s = 0; s < n; ++s) {
            const enbl_stack& S = st[s];
            const StackLayout& Y = lay[(size_t)s];
            if (l + 1 >= S.levels) continue;
            const Level& D = Y.img[(size_t)l];          // image level l + 1
            const Level& DA = Y.ial[(size_t)l];
            const bool keep_alpha = l + 2 < S.levels;   // level l + 1 is reduced again
            const int sw = l == 0 ? S.w : Y.img[(size_t)l - 1].w, sh = l == 0 ? S.h : Y.img[(size_t)l - 1].h;
            auto add_job = [&](int kind, const void* src, long spitch, long splane, int sw_, int sh_, const uint8_t* sa,
                               long sap, void* dst, const Level& DL, uint8_t* da, long dap) {
                jobs.emplace_back();
                BReduceJob& J = jobs.back();
                ReduceArgs& a = J.a;
                a.src = src; a.spitch = spitch; a.splane = splane; a.sw = sw_; a.sh = sh_;
                a.salpha = sa; a.sapitch = sap;
                a.dst = dst; a.dpitch = DL.pitch; a.dplane = DL.plane; a.dw = DL.w; a.dh = DL.h;
                a.dalpha = da; a.dapitch = dap;
                a.channels = kind == kReducePlane1 ? 1 : 3;
                a.wrap = 0;
                a.rows = tu.rows[kBatchReduce];
                a.force_small = ENBL_KERNELS_STREAMING;
                a.frac = 7;
                J.kind = kind;
                J.groups_x = ((DL.w + kStrip - 1) / kStrip + 3) / 4;
                t0.push_back(total);
                total += J.groups_x * ((DL.h + a.rows - 1) / a.rows);
            };
            for (int m = 0; m < S.K; ++m) {
                uint8_t* da = keep_alpha ? (uint8_t*)at(Y.ial_at(S.K, m) + DA.offset) : nullptr;
                if (l == 0) {
                    const bool wide = itype == ENBL_U16 && ((uintptr_t)S.imgs[m] & 3) == 0 && (S.ipitches[m] & 1) == 0 &&
                                      ((uintptr_t)S.alphas[m] & 1) == 0 && (S.apitches[m] & 1) == 0;
                    add_job(wide ? kReduceImage3W : kReduceImage3, S.imgs[m], S.ipitches[m], 0, sw, sh, S.alphas[m],
                            S.apitches[m], at(Y.img_at(m) + D.offset), D, da, DA.pitch);
                } else {
                    const Level& SL = Y.img[(size_t)l - 1];
                    const Level& SA = Y.ial[(size_t)l - 1];
                    add_job(kReducePlane3, at(Y.img_at(m) + SL.offset), SL.pitch, SL.plane, sw, sh,
                            (const uint8_t*)at(Y.ial_at(S.K, m) + SA.offset), SA.pitch, at(Y.img_at(m) + D.offset), D, da,
                            DA.pitch);
                }
                bytes += (double)sw * sh * (3.0 * (l == 0 ? isz : es) + 1.0) + (double)D.w * D.h * (3.0 * es + (keep_alpha ? 1.0 : 0.0));
            }
            const Level& MS = Y.msk[(size_t)l];
            const Level& MD = Y.msk[(size_t)l + 1];
            const Level& UA = Y.ual[(size_t)l];
            const Level& UD = Y.ual[(size_t)l + 1];
            for (int m = 0; m < S.K; ++m) {
                const bool write_ua = m == 0 && keep_alpha;
                add_job(kReducePlane1, at(Y.msk_at(S.K, m) + MS.offset), MS.pitch, MS.plane, MS.w, MS.h,
                        (const uint8_t*)at(Y.ual_at(S.K) + UA.offset), UA.pitch, at(Y.msk_at(S.K, m) + MD.offset), MD,
                        write_ua ? (uint8_t*)at(Y.ual_at(S.K) + UD.offset) : nullptr, UD.pitch);
                bytes += (double)MS.w * MS.h * (ms + (m == 0 ? 1.0 : 0.0)) + (double)MD.w * MD.h * (ms + (write_ua ? 1.0 : 0.0));
            }
        }
        add_launch(kBatchReduce, 0, T.put(jobs.data(), jobs.size()), t0, total, bytes);
    }

    // Merge + collapse passes, coarsest level first.
    for (int l = Lmax - 1; l >= 0; --l) {
        std::vector<BFuseJob> jobs;
        std::vector<int> t0;
        int total = 0;
        double bytes = 0;
        for (int s = 0; s < n; ++s) {
            const enbl_stack& S = st[s];
            const StackLayout& Y = lay[(size_t)s];
            if (l >= S.levels) continue;
            const bool coarsest = l == S.levels - 1;
            BFuseJob J{};
            J.plane0 = (int)planes.size();
            for (int m = 0; m < S.K; ++m) {
                BPlane f, c, k;
                if (l == 0) f = {S.imgs[m], S.ipitches[m], 0};
                else {
                    const Level& F = Y.img[(size_t)l - 1];
                    f = {at(Y.img_at(m) + F.offset), F.pitch, F.plane};
                }
                if (!coarsest) {
                    const Level& C = Y.img[(size_t)l];
                    c = {at(Y.img_at(m) + C.offset), C.pitch, C.plane};
                } else {
                    c = {nullptr, 0, 0};
                }
                const Level& M = Y.msk[(size_t)l];
                k = {at(Y.msk_at(S.K, m) + M.offset), M.pitch, M.plane};
                planes.push_back(f);
                planes.push_back(c);
                planes.push_back(k);
            }
            J.fw = Y.msk[(size_t)l].w;
            J.fh = Y.msk[(size_t)l].h;
            J.cw = coarsest ? 0 : Y.msk[(size_t)l + 1].w;
            J.ch = coarsest ? 0 : Y.msk[(size_t)l + 1].h;
            J.K = S.K;
            J.rcoarse = nullptr;
            J.rcpitch = J.rcplane = 0;
            if (!coarsest) {
                const Level& RC = Y.res[(size_t)l];
                J.rcoarse = at(Y.res_at(S.K) + RC.offset);
                J.rcpitch = RC.pitch;
                J.rcplane = RC.plane;
            }
            J.res = nullptr;
            J.rpitch = J.rplane = 0;
            if (l > 0) {
                const Level& R = Y.res[(size_t)l - 1];
                J.res = at(Y.res_at(S.K) + R.offset);
                J.rpitch = R.pitch;
                J.rplane = R.plane;
            }
            J.out = S.out;
            J.opitch = S.opitch;
            J.ualpha = (const uint8_t*)at(Y.ual_at(S.K) + Y.ual[0].offset);
            J.upitch = Y.ual[0].pitch;
            J.oalpha = S.out_alpha;
            J.oapitch = S.oapitch;
            J.seed = S.seed;
            int nt;
            if (coarsest) {
                nt = tiles(J.fw, J.fh, tu.rows[kBatchFuse], &J.tiles_x);
            } else {
                J.tiles_x = ((J.cw + kStrip - 1) / kStrip + 3) / 4;
                nt = J.tiles_x * ((J.ch + tu.rows[kBatchFuse] - 1) / tu.rows[kBatchFuse]);
            }
            jobs.push_back(J);
            t0.push_back(total);
            total += nt;
            const double fpx = (double)J.fw * J.fh, cpx = (double)J.cw * J.ch;
            bytes += (double)S.K * (fpx * (3.0 * (l == 0 ? isz : es) + ms) + cpx * 3.0 * es) + cpx * 3.0 * es +
                     (l == 0 ? fpx * (3.0 * isz + 2.0) : fpx * 3.0 * es);
        }
        add_launch(kBatchFuse, l == 0, T.put(jobs.data(), jobs.size()), t0, total, bytes);
    }
    P.images_off = T.put(images.data(), images.size());
    P.planes_off = T.put(planes.data(), planes.size());
}

int run_launches(enbl_context* ctx, int itype, const Plan& P, char* dtab, int time_kind, float* kind_ms)
{
    const hipStream_t s = ctx->stream;
    std::vector<hipEvent_t> ev;
    int rc = ENBL_OK;
    for (const Launch& L : P.launches) {
        if (L.kind == kBatchWeights && P.stack_weights) continue;
        const bool timed = kind_ms && L.kind == time_kind;
        if (timed) {
            hipEvent_t a, b;
            rc = hip_check(hipEventCreate(&a), "event");
            if (rc) break;
            rc = hip_check(hipEventCreate(&b), "event");
            if (rc) { (void)hipEventDestroy(a); break; }
            ev.push_back(a);
            ev.push_back(b);
            (void)hipEventRecord(a, s);
        }
        {
            Timed t(ctx, kKindName[L.kind], L.bytes);
            const int* t0 = reinterpret_cast<const int*>(dtab + L.tile0);
            if (L.kind == kBatchWeights)
                rc = launch_batch_weights(itype, P.max_k, reinterpret_cast<const BWeightJob*>(dtab + L.jobs), t0, L.njobs, L.total,
                                          reinterpret_cast<const BImage*>(dtab + P.images_off),
                                          reinterpret_cast<const float*>(dtab + P.lut_off), P.w_exp, P.w_sat, P.scale,
                                          L.rows, L.grid, s);
            else if (L.kind == kBatchReduce)
                rc = launch_batch_reduce(itype, reinterpret_cast<const BReduceJob*>(dtab + L.jobs), t0, L.njobs, L.total,
                                         L.grid, s);
            else
                rc = launch_batch_fuse(itype, P.frac, L.level0, reinterpret_cast<const BFuseJob*>(dtab + L.jobs), t0,
                                       L.njobs, L.total, reinterpret_cast<const BPlane*>(dtab + P.planes_off), P.maxpyr, P.dither, L.rows, L.grid, s);
            if (!rc) rc = launch_status(kKindName[L.kind]);
        }
        if (timed) (void)hipEventRecord(ev.back(), s);
        if (rc) break;
    }
    if (kind_ms) {
        float total = 0.f;
        if (!rc) rc = hip_check(hipStreamSynchronize(s), "batch timing");
        for (size_t i = 0; i + 1 < ev.size(); i += 2) {
            float ms = 0.f;
            if (!rc) rc = hip_check(hipEventElapsedTime(&ms, ev[i], ev[i + 1]), "batch timing");
            total += ms;
        }
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        *kind_ms = total;
    }
    return rc;
}

// Plans the batch, sizes and fills the device arena and table buffer.
int prepare(enbl_context* ctx, BatchState* b, int itype, int n, const enbl_stack* st, const enbl_enfuse_params* p,
            int dither, const Tuning& tu, Plan& P, bool* same, bool stack_weights = false)
{
    P.stack_weights = stack_weights;
    std::vector<StackLayout> lay;
    size_t arena_need = 0;
    layout_stacks(itype, n, st, lay, &arena_need);
    int rc = grow(ctx, b, &b->arena, &b->arena_bytes, std::max<size_t>(arena_need, 256), "batch arena");
    if (rc) return rc;
    static const std::vector<float> no_lut;
    const std::vector<float>& lut = stack_weights ? no_lut : lut_for(b, itype, p);
    const int mfrac = itype == ENBL_U8 ? 7 : 15;
    P.w_exp = p->w_exposure;
    P.w_sat = p->w_saturation;
    P.scale = (double)(1u << mfrac);
    P.maxpyr = (double)(255 << mfrac);
    P.frac = 7;
    P.dither = dither;
    build_plan(itype, n, st, lay, (char*)b->arena, nullptr, lut, tu, b->cus, P);
    rc = grow(ctx, b, &b->tables, &b->table_bytes, P.tab.size(), "batch tables");
    if (rc) return rc;
    build_plan(itype, n, st, lay, (char*)b->arena, (char*)b->tables, lut, tu, b->cus, P);
    // Key of the captured graph: table image, launch list and scalar arguments.
    std::vector<unsigned char> key(P.tab);
    auto append = [&](const void* d, size_t k) { key.insert(key.end(), (const unsigned char*)d, (const unsigned char*)d + k); };
    for (const Launch& L : P.launches) append(&L, sizeof L);
    const double sc[4] = {P.w_exp, P.w_sat, P.scale, P.maxpyr};
    append(sc, sizeof sc);
    const int iv[5] = {P.frac, P.dither, itype, P.max_k, P.stack_weights ? 1 : 0};
    append(iv, sizeof iv);
    *same = b->exec && b->exec_stream == ctx->stream && key == b->key;
    if (*same) return ENBL_OK;
    drop_graph(b);
    rc = hip_check(hipMemcpyAsync(b->tables, P.tab.data(), P.tab.size(), hipMemcpyHostToDevice, ctx->stream), "batch tables");
    if (!rc) rc = hip_check(hipStreamSynchronize(ctx->stream), "batch tables");
    if (rc) return rc;
    b->key.swap(key);   // valid once a graph is instantiated for it
    return ENBL_OK;
}

int run_batch(enbl_context* ctx, int itype, int n, const enbl_stack* st, const enbl_enfuse_params* p, int dither,
              bool stack_weights = false)
{
    BatchState* b;
    int rc = state(ctx, &b);
    if (rc) return rc;
    Plan P;
    bool same = false;
    rc = prepare(ctx, b, itype, n, st, p, dither, context_tuning(ctx), P, &same, stack_weights);
    if (rc) return rc;
    const hipStream_t s = ctx->stream;
    if (stack_weights) {
        // the weights the batched pass does not cover, stack by stack, into
        // the arena's mask levels 0 and union alphas (ahead of the batched
        // passes on the same stream)
        for (int k = 0; k < n && !rc; ++k) {
            const enbl_stack& S = st[k];
            rc = enfuse_stack_masks(ctx, itype, S.w, S.h, S.K, S.imgs, S.ipitches, S.alphas, S.apitches, p,
                                    P.ualpha[(size_t)k], P.upitch[(size_t)k], P.mask0[(size_t)k].data(),
                                    P.mpitch[(size_t)k]);
        }
        if (rc) return rc;
    }
    // Profiling times every launch with events: direct launches then.
    const bool graph = !ctx->profile && s != nullptr;
    if (same && graph) return hip_check(hipGraphLaunch(b->exec, s), "batch graph launch");
    if (same) return run_launches(ctx, itype, P, (char*)b->tables, -1, nullptr);
    if (graph && hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        rc = run_launches(ctx, itype, P, (char*)b->tables, -1, nullptr);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (!rc) rc = hip_check(e, "batch graph capture");
        if (!rc) rc = hip_check(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "batch graph instantiate");
        if (g) (void)hipGraphDestroy(g);
        if (rc) { drop_graph(b); return rc; }
        b->exec_stream = s;
        return hip_check(hipGraphLaunch(b->exec, s), "batch graph launch");
    }
    (void)hipGetLastError();
    b->key.clear();
    return run_launches(ctx, itype, P, (char*)b->tables, -1, nullptr);
}

}  // namespace

extern "C" {

int enbl_enfuse_batch(enbl_context* ctx, int itype, int nstacks, const enbl_stack* stacks,
                      const enbl_enfuse_params* p, int wrap, int dither)
{
    if (!ctx || !p) return fail(ENBL_E_ARG, "enbl_enfuse_batch: null argument");
    int rc = check_stacks(itype, nstacks, stacks, p);
    if (rc || nstacks == 0) return rc;
    // 8/16-bit and float stacks take the batched pyramid passes, with the
    // batched weights pass when it covers the weights (integer images only)
    // and stack-by-stack weights otherwise.  A blend colour space (which
    // converts every image around the pyramids) and wraparound run through
    // enbl_enfuse one stack at a time.
    if (ctx->colorspace == ENBL_COLORSPACE_IDENTITY && batchable(itype, p, wrap))
        return run_batch(ctx, itype, nstacks, stacks, p, dither);
    if (ctx->colorspace == ENBL_COLORSPACE_IDENTITY && batchable_pyramids(itype, wrap))
        return run_batch(ctx, itype, nstacks, stacks, p, dither, true);
    for (int s = 0; s < nstacks; ++s) {
        const enbl_stack& S = stacks[s];
        rc = enbl_enfuse(ctx, itype, S.w, S.h, S.K, S.imgs, S.ipitches, S.alphas, S.apitches, p, S.levels, wrap,
                         S.out, S.opitch, S.out_alpha, S.oapitch, dither, S.seed);
        if (rc) return rc;
    }
    return ENBL_OK;
}

int enbl_batch_get_tuning(const enbl_context* ctx, enbl_batch_tuning* t)
{
    if (!ctx || !t) return fail(ENBL_E_ARG, "enbl_batch_get_tuning: null argument");
    const Tuning tu = context_tuning(ctx);
    for (int k = 0; k < kBatchKinds; ++k) {
        t->rows[k] = tu.rows[k];
        t->blocks_per_cu[k] = tu.bpc[k];
    }
    return ENBL_OK;
}

int enbl_batch_set_tuning(enbl_context* ctx, const enbl_batch_tuning* t)
{
    if (!ctx || !t) return fail(ENBL_E_ARG, "enbl_batch_set_tuning: null argument");
    for (int k = 0; k < kBatchKinds; ++k) {
        if (t->rows[k] < 0 || t->rows[k] > 256 || (t->rows[k] & 3))
            return fail(ENBL_E_ARG, "enbl_batch_set_tuning: rows must be a multiple of 4 in [0, 256]");
        if (t->blocks_per_cu[k] < 0 || t->blocks_per_cu[k] > 64)
            return fail(ENBL_E_ARG, "enbl_batch_set_tuning: blocks per CU must be in [0, 64]");
    }
    for (int k = 0; k < kBatchKinds; ++k) {
        ctx->batch_rows[k] = t->rows[k];
        ctx->batch_bpc[k] = t->blocks_per_cu[k];
    }
    return ENBL_OK;
}

int enbl_batch_autotune(enbl_context* ctx, int itype, int nstacks, const enbl_stack* stacks,
                        const enbl_enfuse_params* p, int dither, int repeats, enbl_batch_tuning* best)
{
    if (!ctx || !p || !best) return fail(ENBL_E_ARG, "enbl_batch_autotune: null argument");
    int rc = check_stacks(itype, nstacks, stacks, p);
    if (rc) return rc;
    if (nstacks == 0 || !batchable(itype, p, 0) || ctx->colorspace != ENBL_COLORSPACE_IDENTITY)
        return fail(ENBL_E_UNSUPPORTED, "enbl_batch_autotune: this batch does not take the batched schedule");
    if (repeats < 1) repeats = 1;
    BatchState* b;
    rc = state(ctx, &b);
    if (rc) return rc;
    // Candidate tile rows per pass kind (the streaming reduce amortises its
    // 2-row halo over longer tiles) and resident blocks per CU.
    static const int kRows[kBatchKinds][4] = {{4, 8, 16, 32}, {8, 16, 32, 64}, {4, 8, 16, 32}};
    static const int kBpc[] = {2, 4, 8, 16};
    Tuning cur = context_tuning(ctx);
    const bool prof = ctx->profile;
    ctx->profile = false;
    for (int kind = 0; kind < kBatchKinds && !rc; ++kind) {
        float best_ms = 1e30f;
        Tuning win = cur;
        for (int r : kRows[kind]) {
            for (int q : kBpc) {
                Tuning t = cur;
                t.rows[kind] = r;
                t.bpc[kind] = q;
                Plan P;
                bool same;
                rc = prepare(ctx, b, itype, nstacks, stacks, p, dither, t, P, &same);
                if (rc) break;
                b->key.clear();   // direct launches below; no graph stands for this table
                float ms = 1e30f;
                for (int i = 0; i <= repeats && !rc; ++i) {   // first run warms up
                    float k_ms = 0.f;
                    rc = run_launches(ctx, itype, P, (char*)b->tables, kind, &k_ms);
                    if (i > 0) ms = std::min(ms, k_ms);
                }
                if (rc) break;
                if (ms < best_ms) {
                    best_ms = ms;
                    win = t;
                }
            }
            if (rc) break;
        }
        cur = win;
    }
    ctx->profile = prof;
    if (rc) return rc;
    for (int k = 0; k < kBatchKinds; ++k) {
        ctx->batch_rows[k] = cur.rows[k];
        ctx->batch_bpc[k] = cur.bpc[k];
        best->rows[k] = cur.rows[k];
        best->blocks_per_cu[k] = cur.bpc[k];
    }
    return ENBL_OK;
}

}  // extern "C"
