// The multi-job small-level kernel of the enblend step: one launch runs
// several independent jobs (enbl_tail.h), a block finding its job by the
// jobs' block ranges (read from the kernel-argument table with scalar
// loads).
#include <type_traits>

#include "enbl_tail.h"

namespace enbl {
namespace {

template <class P, class M, int C>
__global__ __launch_bounds__(256) void tail_kernel(TailLaunch t)
{
    const int b = blockIdx.x;
    int k = 0;
    for (int q = 1; q < t.njobs; ++q)
        if (b >= t.job[q].block0) k = q;
    const TailJob& J = t.job[k];
    const int local = b - J.block0;
    const int bx = local % J.tiles_x, by = local / J.tiles_x;
    switch (J.kind) {
    case kTailReduceImage: reduce_small_body<P, C, true, false>(J.u.r, bx, by); break;
    case kTailReduceMask: reduce_small_body<M, 1, false, false>(J.u.r, bx, by); break;
    case kTailReduceMaskAlpha: reduce_small_body<M, 1, true, false>(J.u.r, bx, by); break;
    case kTailExpandSub: expand_small_body<P, C, false, false>(J.u.e, bx, by); break;
    case kTailBlend: blend_small_body<P, M, C>(J.u.b, bx, by); break;
    case kTailCollapse: collapse_small_body<P, M, C>(J.u.c, bx, by); break;
    case kTailReduceImageStrip:
    case kTailReduceMaskStrip:
    case kTailReduceMaskAlphaStrip: {
        const int wx = bx * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const int i0 = by * J.u.r.rows;
        if (J.kind == kTailReduceImageStrip) reduce_strip_ring<P, C, true, PlaneSrc<P, C>, true>(J.u.r, wx, i0);
        else if (J.kind == kTailReduceMaskStrip) reduce_strip_ring<M, 1, false, PlaneSrc<M, 1>, true>(J.u.r, wx, i0);
        else reduce_strip_ring<M, 1, true, PlaneSrc<M, 1>, true>(J.u.r, wx, i0);
        break;
    }
    default: break;
    }
}

template <class P, class M>
void launch_t(int C, const TailLaunch& t, hipStream_t s)
{
    if (C == 1) hipLaunchKernelGGL((tail_kernel<P, M, 1>), dim3(t.blocks), dim3(256), 0, s, t);
    else hipLaunchKernelGGL((tail_kernel<P, M, 3>), dim3(t.blocks), dim3(256), 0, s, t);
}

}  // namespace

int launch_tail(int ptype, int mtype, int C, const TailLaunch& t, hipStream_t s)
{
    if ((C != 1 && C != 3) || t.njobs < 1 || t.njobs > kTailMaxJobs || t.blocks < 1) return ENBL_E_ARG;
    if (ptype == ENBL_I16 && mtype == ENBL_I16) { launch_t<int16_t, int16_t>(C, t, s); return 0; }
    if (ptype == ENBL_I32 && mtype == ENBL_I32) { launch_t<int32_t, int32_t>(C, t, s); return 0; }
    if (ptype == ENBL_F64 && mtype == ENBL_I32) { launch_t<double, int32_t>(C, t, s); return 0; }
    return ENBL_E_UNSUPPORTED;
}

}  // namespace enbl
