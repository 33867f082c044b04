// Host-side internals shared by the launchers and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/enblend_amd.h"

namespace enbl {

inline int ptype_size(int t) {
    switch (t) {
    case ENBL_I16: return 2;
    case ENBL_I32: return 4;
    case ENBL_F64: return 8;
    case ENBL_U8: return 1;
    case ENBL_U16: return 2;
    case ENBL_F32: return 4;
    }
    return 0;
}

struct ReduceArgs {
    const void* src; long spitch; long splane; int sw; int sh;
    const uint8_t* salpha; long sapitch;
    void* dst; long dpitch; long dplane; int dw; int dh;
    uint8_t* dalpha; long dapitch;
    int channels; int wrap; int rows;
    int force_small;   // ENBL_KERNELS_* policy
    int frac;          // interleaved integer image source: fixed-point shift
    int pad_;          // 0 (no uninitialised bytes: batched job tables are compared bytewise)
    // Image sources of the level-0 reduce of a float image: the converted
    // level 0 is also written here (planes, pitch in elements) when non-null.
    void* conv; long cpitch; long cplane;
};

// A mask pyramid's reduce from level 0 formed on the fly: src is the image's
// float weight plane, norm the float norm plane (pitch in elements), and each
// level-0 value fromRealPromote((hard ? w : norm > 0 ? 255 w / norm :
// fallback) * scale).
struct NormReduceArgs : ReduceArgs {
    const float* norm; long npitch;
    double scale;
    float fallback;
    int hard;
};

struct ExpandArgs {
    const void* src; long spitch; long splane; int sw; int sh;
    void* dst; long dpitch; long dplane; int dw; int dh;
    int channels; int wrap; int add; int rows;
    int force_small;
};

int launch_reduce(int ptype, ReduceArgs a, hipStream_t s);
// Streaming reduce whose source is an interleaved integer image of itype
// (level 0 converted on the fly, a.frac); no wraparound.
int launch_reduce_image(int itype, int ptype, ReduceArgs a, hipStream_t s);
int launch_reduce_norm_mask(int mtype, NormReduceArgs a, hipStream_t s);
int launch_expand(int ptype, ExpandArgs a, hipStream_t s);

// Pointwise kernels (pointwise.hip).
int launch_to_pyramid(int itype, int ptype, int frac, int channels,
                      const void* img, long ipitch, int w, int h,
                      void* planes, long ppitch, long pplane, hipStream_t s);
int launch_from_pyramid_if(int ptype, int itype, int frac, int channels,
                           const void* planes, long ppitch, long pplane, int w, int h,
                           const uint8_t* mask, long mpitch, void* img, long ipitch,
                           int dither, uint64_t seed, hipStream_t s);
int launch_blend(int mtype, int ptype, int channels, int w, int h,
                 const void* mask, long mpitch,
                 const void* white, long wpitch, long wplane,
                 void* black, long bpitch, long bplane,
                 double white_value, hipStream_t s);
int launch_alpha_union(const uint8_t* src, long spitch, uint8_t* dst, long dpitch, int w, int h,
                       hipStream_t s);

struct Level {
    int w, h;
    long pitch;   // elements
    long plane;   // elements between channel planes
    size_t offset;  // bytes from the arena base
};

std::vector<Level> plan_levels(int ptype, int channels, int w, int h, int levels, size_t* bytes);

}  // namespace enbl

struct enbl_context;
struct enbl_pyramid;

constexpr int kScratchSlots = 40;

namespace enbl {
struct BatchState;   // batched enfuse: arena, job tables, captured graph (batch_capi.hip)
namespace host {
void batch_release(enbl_context* ctx);
}
}  // namespace enbl

namespace enbl {
namespace host {

// Error reporting: stores the message for enbl_last_error and returns code.
int fail(int code, const char* fmt, ...);
int hip_check(hipError_t e, const char* what);
int launch_status(const char* what);

// Brackets one launch with HIP events when the context profiles; `bytes` is
// the launch's credited algorithmic byte count (the unfused operations it
// performs), `own` the kernel's own minimum traffic (negative: same as bytes).
struct Timed {
    enbl_context* ctx;
    size_t idx;
    Timed(enbl_context* c, const char* name, double bytes, double own = -1.0);
    ~Timed();
};

const char* pname(int t);
size_t round_up(size_t v, size_t m);
int mask_ptype_for(int itype);
int image_ptype_for(int itype);
void* level_ptr(const enbl_pyramid* p, int l);
int gaussian_levels(enbl_context* ctx, enbl_pyramid* p, int wrap, const uint8_t* alpha, long apitch);
// Levels 1.. of a Gaussian pyramid whose level 0 is built, with the per-level
// alpha ping-ponging between a0 and a1 (pitch apl, each at least apl rows of
// level 1; unused without alpha).
int gaussian_levels_with(enbl_context* ctx, enbl_pyramid* p, int wrap, const uint8_t* alpha, long apitch,
                         uint8_t* a0, uint8_t* a1, long apl);
int laplacian_levels(enbl_context* ctx, enbl_pyramid* p, int wrap);
int build_level0(enbl_context* ctx, enbl_pyramid* p, int itype, int frac, const void* img, long ipitch);
// Pipeline scratch pyramids, kept in the context and keyed by role, type,
// channels, geometry and slot, so pyramids of different roles never alias
// even where their type and channel count coincide (grayscale images whose
// pyramid type equals the mask pyramid type).
enum PyramidRole { kRoleImage = 0, kRoleMask = 1, kRoleResult = 2, kRoleAdjoint = 3 };
int cached_pyramid(enbl_context* ctx, int role, int ptype, int C, int w, int h, int levels, int slot,
                   enbl_pyramid** out);
// Device scratch buffer `slot` of the context, grown to at least `bytes`.
int scratch(enbl_context* ctx, int slot, size_t bytes, void** out);
// Whether the pass producing a coarse level of cw x ch takes the streaming
// (and fused) kernels under the current kernel policy.
bool streaming_level(int cw, int ch);
int alpha_pingpong(enbl_context* ctx, const enbl_pyramid* p, uint8_t** a, uint8_t** b, long* pitch);
// Blend colour spaces (colorspace.hip): validity for an image type and
// channel count, and the forward / inverse conversions of RGB images.
int check_blend_space(int space, int itype, int channels);
int rgb_to_blend_space(enbl_context* ctx, int space, int itype, int w, int h, const void* img, long ip, void* out,
                       long op);
int blend_space_to_rgb_if(enbl_context* ctx, int space, int itype, int w, int h, const void* src, long sp,
                          const uint8_t* mask, long mp, void* img, long ip);
// CIECAM02: RGB into the planar JCh pyramid level 0 (pitch op, channel
// planes opl apart), and a collapsed JCh level 0 back to RGB under a mask.
int rgb_to_jch(enbl_context* ctx, int itype, int w, int h, const void* img, long ip, void* out, long op, long opl);
int jch_to_rgb_if(enbl_context* ctx, int itype, int w, int h, const void* lv, long lp, long lpl, const uint8_t* mask,
                  long mp, void* img, long ip);
int laplacian_pyramid_any(enbl_context* ctx, int wrap, int itype, int frac, const void* img, long ipitch,
                          const uint8_t* alpha, long apitch, enbl_pyramid* out);
// enfuse's K Laplacian pyramids (image alphas) and K mask Gaussian pyramids
// (one shared alpha, level 0 written) without wraparound, the small levels
// of each set in shared multi-job launches.
// With g1s (one buffer of level-1 geometry per image): level 0 is only
// reduced, from the image, into g1s[m], and the pyramids hold Laplacian
// levels 1.. and, for float images, the converted level 0 (Gaussian, not
// Laplacian; integer images leave level 0 unwritten).  Needs
// enfuse_level0_count >= 1.
int enfuse_image_pyramids(enbl_context* ctx, int itype, int frac, int K, const void* const* imgs, const long* ips,
                          const uint8_t* const* alphas, const long* aps, enbl_pyramid* const* lps,
                          void* const* g1s = nullptr);
// Levels whose coarse level is built by a streaming pass (no wraparound).
int enfuse_level0_count(const enbl_pyramid* p, const uint8_t* alpha);
// collapsePyramid (src/pyramid.h:1837-1886) of levels levels-2 .. lmin.
int collapse_levels(enbl_context* ctx, int wrap, enbl_pyramid* p, int lmin);
// l0: level 0 of each mask is not stored but formed from the images' weight
// planes and the norm (the normalisation pass's values) by the first reduce
// (needs a streaming level 1).
struct MaskLevel0 {
    const float* const* w; long wpitch;
    const float* norm; long npitch;
    double scale;
    float fallback;
    int hard;
};
int enfuse_mask_pyramids(enbl_context* ctx, int K, enbl_pyramid* const* mgps, const uint8_t* ua, long uap,
                         const MaskLevel0* l0 = nullptr);
// enbl_enfuse's weights of one stack normalised into the fixed-point mask
// level 0 of every image (mask0[m], pitch mpitch in elements) and its union
// alpha (ua, pitch uap).
int enfuse_stack_masks(enbl_context* ctx, int itype, int w, int h, int K, const void* const* imgs, const long* ips,
                       const uint8_t* const* alphas, const long* aps, const enbl_enfuse_params* p, uint8_t* ua,
                       long uap, void* const* mask0, long mpitch);
// Overlapping work: fork_aux makes the auxiliary stream wait for the work
// issued so far and directs later launches to it; end_aux records its end
// and directs launches back to the main stream; join_aux makes the main
// stream wait for the auxiliary work.  sync_streams waits for both.
int fork_aux(enbl_context* ctx);
int end_aux(enbl_context* ctx);
int join_aux(enbl_context* ctx);
int sync_streams(enbl_context* ctx);

}  // namespace host
}  // namespace enbl

struct enbl_context {
    int device;
    hipStream_t own_stream;
    hipStream_t stream;
    // A second stream for work that overlaps the main sequence
    // (fork_aux / end_aux / join_aux); `forked_from` is the main stream while
    // work is being issued on aux_stream, else null.
    hipStream_t aux_stream;
    hipEvent_t aux_fork, aux_join, aux_mid;
    hipStream_t forked_from;
    bool profile;
    struct Timed { std::string name; double bytes, own; hipEvent_t start, stop; };
    std::vector<Timed> timed;
    struct Cached { int role, ptype, channels, w, h, levels; enbl_pyramid* p; };
    std::vector<Cached> cache;   // pipeline scratch pyramids
    void* scratch_alpha;         // per-level alpha scratch for Gaussian builds
    size_t scratch_alpha_bytes;
    void* scratch_buf[kScratchSlots];     // general device scratch (weights, fused passes, NFT)
    size_t scratch_bytes[kScratchSlots];
    enbl::BatchState* batch;     // batched enfuse state, created on first use
    int batch_rows[3];           // launch shape of the batched passes (0: default)
    int batch_bpc[3];
    int colorspace;              // ENBL_COLORSPACE_* of the pipelines (enbl_set_blend_colorspace)
};

struct enbl_pyramid {
    enbl_context* ctx;
    int ptype, channels, w, h, levels;
    void* base;
    size_t bytes;
    std::vector<enbl::Level> lv;
};
