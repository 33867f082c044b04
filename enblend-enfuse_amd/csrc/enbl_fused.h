// Launch arguments of the fused pyramid passes (fused.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace enbl {

struct FusedLapArgs {
    const void* src; long spitch; long splane; int frac;   // fine level: planes, or interleaved image
    const uint8_t* salpha; long sapitch;
    int fw, fh;
    void* lap; long lpitch; long lplane;                    // Laplacian level out
    void* coarse; long cpitch; long cplane;                 // next Gaussian level out
    uint8_t* calpha; long capitch;
    int cw, ch;
    int rows;
};

struct FusedCollapseArgs {
    const void* coarse; long cpitch; long cplane; int cw, ch;   // collapsed level l+1
    void* black; long fpitch; long fplane;                      // black Laplacian level l (in/out)
    const void* white; long wpitch; long wplane;                // white Laplacian level l
    int fw, fh;
    const void* mask; long mpitch; int mfrac;                   // mask level l (planes or 0/255 bytes)
    double white_value;
    double wc_scale;                                            // set by the launcher
    void* img; long ipitch; int ifrac;                          // output image (level 0)
    const uint8_t* walpha; long wapitch;
    uint8_t* balpha; long bapitch;
    int dither; uint64_t seed;
    int rows;
    // Level 0 with both Laplacian levels formed on the fly (otf != 0, integer
    // pyramids): white = wimg << ifrac - expand(wg1), black likewise from the
    // black image `img` (read before it is overwritten) and bg1; `white` and
    // `black` are then unused.
    int otf;
    const void* wimg; long wipitch;
    const void* wg1; long wg1pitch; long wg1plane;
    const void* bg1; long bg1pitch; long bg1plane;
};

// src_kind 0: planes of ptype; ENBL_U8/ENBL_U16/ENBL_F32: interleaved image.
int launch_laplacian_level(int ptype, int src_kind, const FusedLapArgs& a, int C, hipStream_t s);
// out_itype 0: write the collapsed level into `black`; otherwise convert into
// the interleaved image of that type (mask given as 0/255 bytes).
int launch_collapse_level(int ptype, int mtype, int out_itype, bool mask8, const FusedCollapseArgs& a, int C,
                          hipStream_t s);

}  // namespace enbl
