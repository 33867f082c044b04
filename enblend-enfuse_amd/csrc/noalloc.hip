// Allocation-free Gaussian pyramid build (enbl_gaussian_pyramid_noalloc,
// declared in include/enblend_amd.h).
//
// enbl_gaussian_pyramid keeps the per-level alpha of a pyramid with alpha in
// two context buffers that it grows with hipMalloc.  This variant takes them
// from `scratch`, a file-scope device array of kNoAllocScratchBytes that every
// GPU the library runs on holds from the first call until the process ends,
// so the call itself allocates no device memory: for callers under HIP graph
// capture or in loops that must not allocate.
//
// In host code `scratch` is only a host-side stand-in for the symbol; its
// device address differs per GPU and is looked up once per device with
// hipGetSymbolAddress (with that device current) and kept in
// g_scratch_ptr[].  Kernels only ever see that address.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (two streams, two threads) race on it.  The reduce passes write every
// alpha value they later read, so nothing in the array needs clearing and the
// contents a previous call left are never read.
#include <mutex>

#include "enbl_internal.h"

using namespace enbl;
using namespace enbl::host;

namespace {

constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;   // 16 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on ctx's GPU, looked up on the first call for
// that GPU with the GPU current.
int scratch_address(const enbl_context* ctx, uint8_t** out)
{
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return fail(ENBL_E_UNSUPPORTED, "enbl_gaussian_pyramid_noalloc: device %d beyond %d", ctx->device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        int prev = -1;
        int rc = hip_check(hipGetDevice(&prev), "hipGetDevice");
        if (rc) return rc;
        if (prev != ctx->device) {
            rc = hip_check(hipSetDevice(ctx->device), "hipSetDevice");
            if (rc) return rc;
        }
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (prev != ctx->device) (void)hipSetDevice(prev);
        rc = hip_check(e, "hipGetSymbolAddress(scratch)");
        if (rc) return rc;
        if (!ptr) return fail(ENBL_E_HIP, "hipGetSymbolAddress(scratch): null address");
        g_scratch_ptr[ctx->device] = ptr;
    }
    *out = (uint8_t*)g_scratch_ptr[ctx->device];
    return ENBL_OK;
}

}  // namespace

extern "C" {

int enbl_gaussian_pyramid_noalloc(enbl_context* ctx, int wrap, int itype, int frac, const void* img, long ipitch,
                                  const uint8_t* alpha, long apitch, enbl_pyramid* out)
{
    if (!ctx || !img || !out) return fail(ENBL_E_ARG, "enbl_gaussian_pyramid_noalloc: null argument");
    uint8_t* a0 = nullptr;
    uint8_t* a1 = nullptr;
    long apl = 0;
    if (alpha && out->levels > 1) {
        apl = (long)round_up((size_t)out->lv[1].w, 256);
        const size_t need = 2 * (size_t)apl * (size_t)out->lv[1].h;
        if (need > kNoAllocScratchBytes)
            return fail(ENBL_E_SIZE,
                        "enbl_gaussian_pyramid_noalloc: level-1 alpha of %dx%d needs %zu bytes of scratch, "
                        "more than the %zu the library holds (use enbl_gaussian_pyramid)",
                        out->lv[1].w, out->lv[1].h, need, kNoAllocScratchBytes);
        uint8_t* base = nullptr;
        int rc = scratch_address(ctx, &base);
        if (rc) return rc;
        a0 = base;
        a1 = base + (size_t)apl * out->lv[1].h;
    }
    // Profiling records HIP events per launch; that is off for this call.
    const bool prof = ctx->profile;
    ctx->profile = false;
    int rc = build_level0(ctx, out, itype, frac, img, ipitch);
    if (!rc) rc = gaussian_levels_with(ctx, out, wrap, alpha, apitch, a0, a1, apl);
    ctx->profile = prof;
    return rc;
}

}  // extern "C"
