// Multi-job launches for the small levels of the enblend step and of
// enfuse's pyramids (tail.hip): reduces, expand-subtracts, blends and blend
// + collapse steps of several pyramids that do not depend on each other run
// in one launch, each job on its own range of 256-thread blocks (64 columns
// x 4 rows of samples).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "enbl_internal.h"
#include "reduce_stream.h"
#include "small_levels.h"

namespace enbl {

enum TailKind : int {
    kTailReduceImage = 0,   // image level with alpha (ReduceArgs, P, C)
    kTailReduceMask = 1,    // mask level without alpha (ReduceArgs, M, 1 channel)
    kTailExpandSub = 2,     // Laplacian level: G_l - expand(G_l+1) (ExpandArgs, P, C)
    kTailBlend = 3,         // blend of one level (BlendArgs)
    kTailCollapse = 4,      // blend of level l + expand of collapsed level l+1 (CollapseSmallArgs)
    kTailReduceImageStrip = 5,   // as kTailReduceImage in streaming strips: 4 strips x r.rows rows per block
    kTailReduceMaskStrip = 6,    // as kTailReduceMask in streaming strips
    kTailReduceMaskAlpha = 7,    // mask level with alpha (ReduceArgs, M, 1 channel; dalpha may be null)
    kTailReduceMaskAlphaStrip = 8,   // as kTailReduceMaskAlpha in streaming strips
};

struct TailJob {
    int kind;
    int tiles_x;   // 64-column tiles (strips: groups of 4 strips) across the job's sample grid
    int block0;    // first block of the job in its launch
    int pad_;
    union {
        ReduceArgs r;
        ExpandArgs e;
        BlendArgs b;
        CollapseSmallArgs c;
    } u;
};

constexpr int kTailMaxJobs = 16;

struct TailLaunch {
    int njobs;
    int blocks;    // total blocks of the launch
    int pad_[2];
    TailJob job[kTailMaxJobs];
};

// ptype / mtype: image and mask pyramid element types; C image channels.
int launch_tail(int ptype, int mtype, int C, const TailLaunch& t, hipStream_t s);

}  // namespace enbl
