// Job descriptors of the batched enfuse passes (batch.hip).
//
// Every pass is one persistent launch over a packed tile list: job j owns
// tiles [tile0[j], tile0[j+1]) of the list, a tile covering 64 columns x
// `rows` rows of the job's work grid.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "enbl_internal.h"

namespace enbl {

struct BPlane {           // planar level (pyramid type), or an interleaved RGB image
    const void* p;
    long pitch;           // elements
    long plane;           // elements between channel planes (unused for an image)
};

// Per-image inputs of the weights pass (one table for the whole launch; a
// job's images are entries img0 .. img0 + K - 1).
struct BImage {
    const void* img; long ipitch;       // interleaved image, pitch in elements
    const uint8_t* alpha; long apitch;  // pitch in bytes
    void* mask0;                        // mask level 0 (pitch: the job's mpitch)
    long pad_;                          // 0
};

// Weights of all K images of a stack, their norm, the union alpha and the
// normalised fixed-point mask level 0 of every image; a tile is 128 columns
// (a pixel pair per lane) x `rows` rows.
struct BWeightJob {
    int w, h, K;
    int tiles_x;
    int img0;                       // first BImage entry of the stack
    int pad0_;                      // 0 (no uninitialised bytes in a job table)
    long mpitch;
    uint8_t* ualpha; long upitch;   // union alpha level 0
    float fallback;                 // (float)(255 / K)
    int pad_;                       // 0
};

// Reduce with alpha of one pyramid level (image or mask pyramid), as the
// streaming strips of reduce_stream.h: a tile is 4 adjacent 62-column wave
// strips x a.rows output rows.
// kReduceImage3W: the interleaved image and its alpha rows aligned for
// whole-word pair loads (4-byte image rows, 2-byte alpha rows).
enum { kReducePlane3 = 0, kReduceImage3 = 1, kReducePlane1 = 2, kReduceImage3W = 3 };
struct BReduceJob {
    ReduceArgs a;                   // no wraparound; a.rows = tile rows; a.dalpha may be null
    int groups_x;                   // 4-strip groups across the destination level
    int kind;                       // kReduce*
};

// Merge of level l fused with the collapse step of level l, one tile = 4
// adjacent wave strips of 62 coarse columns (2 x 2 fine samples each) x
// `rows` coarse rows:
//   acc = sum over images m (in order) of
//         fromRealPromote((G_m[l] - expand(G_m[l+1])) * mask_m[l] / maxpyr)
//   res[l] = sat(acc + expand(res[l+1]))
// At the coarsest level of a stack (cw == 0) the Laplacian level is G_m[l]
// and nothing is added.  At level 0 the value goes straight into the output
// image under the union alpha.
struct BFuseJob {
    int fw, fh, cw, ch;
    int K;
    int tiles_x;                    // tiles across: 4-strip groups, or 64-column tiles at the coarsest level
    int plane0;                     // BPlane entries plane0 + 3 m + {0: fine level l (level 0: the
    int pad_;                       //   image), 1: coarse level l + 1, 2: mask level l} of image m; 0
    const void* rcoarse; long rcpitch; long rcplane;   // result level l+1
    void* res; long rpitch; long rplane;               // result level l (l > 0)
    void* out; long opitch;                            // level 0: output image (elements)
    const uint8_t* ualpha; long upitch;                // level 0: union alpha
    uint8_t* oalpha; long oapitch;                     // level 0: output alpha
    uint64_t seed;                                     // level 0: dither seed
};

enum { kBatchWeights = 0, kBatchReduce = 1, kBatchFuse = 2, kBatchKinds = 3 };

// max_k: the deepest stack of the launch (selects the register-resident instance).
int launch_batch_weights(int itype, int max_k, const BWeightJob* jobs, const int* tile0, int njobs, int total_tiles,
                         const BImage* images, const float* lut, double w_exposure, double w_saturation, double scale, int rows,
                         int grid, hipStream_t s);
int launch_batch_reduce(int itype, const BReduceJob* jobs, const int* tile0, int njobs, int total_tiles, int grid,
                        hipStream_t s);
int launch_batch_fuse(int itype, int frac, int level0, const BFuseJob* jobs, const int* tile0, int njobs,
                      int total_tiles, const BPlane* planes, double maxpyr, int dither, int rows, int grid, hipStream_t s);

// Level 0 of one stack's merge and collapse with every image's Laplacian
// level 0 formed on the fly from the image (ConvertScalarToPyramidFunctor or
// the float log transform) and its Gaussian level 1: the result, plus the
// expanded collapsed result level 1, goes straight into the output image
// under the union alpha, and the union alpha into the output alpha.  Passed
// by value (no tables).
constexpr int kL0Max = 16;
struct L0FuseArgs {
    int K, fw, fh, cw, ch, rows;
    const void* img[kL0Max]; long ipitch[kL0Max];   // interleaved RGB images, pitch in elements
    const void* g0[kL0Max]; long g0pitch, g0plane;   // float images: their converted level 0 (or null)
    const void* g1[kL0Max]; long gpitch, gplane;     // Gaussian level 1 of every image
    const float* wt[kL0Max]; long wpitch;            // weight plane of every image (mask level 0
    const float* norm; long npitch;                  //   = fromRealPromote(normalised weight * mscale),
    double mscale; float fallback; int hard;         //   formed on the fly)
    const void* rc; long rcpitch, rcplane;           // collapsed result level 1
    const uint8_t* ualpha; long upitch;
    void* out; long opitch;                          // pitch in elements
    uint8_t* oalpha; long oapitch;
    double maxpyr;
    int frac, dither;
    uint64_t seed;
};
int launch_enfuse_level0(int itype, L0FuseArgs a, hipStream_t s);

}  // namespace enbl
